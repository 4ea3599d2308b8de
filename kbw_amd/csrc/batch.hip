// Batches of circuits in one persistent launch (kbwh_batch_*).
//
// Each circuit of a batch owns a slice of one HBM buffer (2^max(n, TL)
// amplitudes; a circuit narrower than a tile runs with idle extra qubits that
// stay |0>, which leaves its amplitudes unchanged).  Its gates are planned
// once on the host into tiled passes (the same planner and gate application
// as KBWH_RUN_TILED); its Hamiltonian is evaluated inside passes too:
//   * every diagonal string, and every string whose X/Y qubits are tile
//     qubits of the circuit's last gate pass, is summed from that pass's
//     tiles while they are still in registers;
//   * the other strings get read-only passes whose tiles are chosen to hold
//     their X/Y qubits (greedily, several strings per pass).
// A run is three kernels:
//   k_coef   -- the coefficients of every parameterised gate from the
//               parameter vector (device sincos), once per run;
//   k_reset  -- the work counter back to 0 (a kernel, not a memset, so the
//               run can be captured in a hipGraph);
//   k_batch  -- persistent: every workgroup takes work items from the
//               counter, largest circuits first.  An item is a chunk of a
//               circuit's tiles (a circuit of up to 32 tiles is one chunk,
//               larger ones up to 16 chunks of about 32 tiles); the item runs
//               every pass of its circuit over its tiles (the first pass
//               makes |0...0> in registers instead of loading), and the
//               chunks of one circuit meet at a barrier between passes.  A
//               workgroup exits when the counter runs past the last item, so
//               every wave reaches the exit.
// With graphs enabled, the parameter upload, the three kernels and the result
// download are captured once and replayed by later runs.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "internal.h"
#include "state.h"
// batch energies agree with the reference to rounding (angles come from the
// device's sincos), so gate products may fuse into FMAs
#define KBWH_TILE_CONTRACT 1
#include "tile_core.h"

#pragma clang fp contract(off)

using namespace kbwh;
using namespace kbwh::tile;

namespace {

constexpr uint32_t kMaxBatchQubits = 26;
constexpr uint32_t kPassInit = 1;   // first pass: |0...0> instead of a load
constexpr uint32_t kPassStore = 2;  // gate pass: write the tile back
constexpr uint32_t kPassLds = 4;    // some string of the pass has X/Y qubits

struct BCirc {
    uint64_t state_off;   // first amplitude of the circuit's slice
    uint32_t n_eff;       // max(n, TL)
    uint32_t first_pass, num_passes;
    uint32_t num_strings;
    uint32_t pad[2];
};

// One work item: tiles [first_tile, first_tile + num_tiles) of circuit circ
// in every pass; the circuit's chunks are consecutive items.
struct BItem {
    uint32_t circ, chunk, nchunks;
    uint32_t first_tile, num_tiles;
    uint32_t pad[3];
};

// Chunks: a circuit of more than kChunkTiles tiles is split into up to
// kMaxChunks chunks.  The persistent grid must hold at least kMaxChunks
// resident workgroups (it holds 4 per CU), so the chunks of a circuit that
// has started are running or are the next items a free workgroup takes.
constexpr uint32_t kChunkTiles = 32;
constexpr uint32_t kMaxChunks = 16;
// A barrier wait gives up after this many polls (about a second): the
// workgroup then goes on, the error flag is set and the host reports it, so
// the kernel ends even if a chunk never arrived.
constexpr uint32_t kSpinLimit = 1u << 24;

struct BPass {
    TileSpec ts;
    uint32_t first_gate, num_gates;
    uint32_t first_str, num_str;
    uint32_t flags, pad;
};

// P|j> = i^ny (-1)^popcount(j & z) |j ^ x>; xt is x as a mask of tile positions
struct BString {
    uint64_t x, z;
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
    return t;  // the same value in every thread
}

// coefficients of parameterised gates: the template's pad field holds the
// absolute parameter index (or -1), the live copy gets the coefficients
__global__ __launch_bounds__(256) void k_coef(const TGate *__restrict__ tmpl, TGate *__restrict__ live,
                                              uint64_t ngates, const double *__restrict__ params) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ngates) return;
    TGate g = tmpl[i];
    const int32_t p = (int32_t)g.pad;
    if (p >= 0) set_coefficients_dev(g, params[p]);
    live[i] = g;
}

// Zeroes the control words before a run (a kernel, not a memset, so the run
// can be captured in a hipGraph): [0] item counter, [1] error flag, then a
// barrier counter and a done counter per circuit.
__global__ __launch_bounds__(256) void k_reset(uint32_t *ctl, uint32_t count) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) ctl[i] = 0;
}

// All chunks of a circuit have finished pass p before any starts pass p + 1:
// the workgroup's stores are made visible device-wide, thread 0 arrives and
// polls until every chunk of the circuit has arrived for this pass (target =
// (p + 1) * chunks), and the wait ends with an acquire, so the next pass
// reads what the other chunks wrote.
__device__ __forceinline__ void circuit_barrier(uint32_t *bar, uint32_t target, uint32_t *err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        uint32_t spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(4);
            if (++spins >= kSpinLimit) {
                atomicOr(err, 1u);
                break;
            }
        }
        __threadfence();
    }
    __syncthreads();
}

__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void k_batch(
    dv2 *__restrict__ states, const BCirc *__restrict__ circs, const BItem *__restrict__ items,
    uint32_t nitems, const BPass *__restrict__ passes, const TGate *__restrict__ gates,
    const BString *__restrict__ strings, double *__restrict__ results, double *__restrict__ partial,
    uint32_t *__restrict__ ctl, uint32_t ncirc) {
    uint32_t *const counter = ctl, *const err = ctl + 1, *const bar = ctl + 2, *const done = ctl + 2 + ncirc;
    __shared__ dv2 lds[TT * TR];  // gate exchanges use the first TT * kChunk slots
    __shared__ double ws[TT / 64];
    __shared__ uint32_t s_item;
    const uint32_t tid = threadIdx.x;
    // Thread 0 fetches the next item in one place per iteration (together
    // with the result store) and every wave reads it back as a uniform value:
    // with a second `tid == 0` block at the loop top the compiler merged the
    // two into a loop nest in which lanes 1..63 re-entered the item read
    // without thread 0 fetching, so the item never advanced.
    if (tid == 0) s_item = atomicAdd(counter, 1u);
    __syncthreads();
    for (;;) {
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= nitems) break;
        const BItem I = items[item];
        const uint32_t ci = I.circ;
        const BCirc C = circs[ci];
        dv2 *a = states + C.state_off;
        const uint32_t t_end = I.first_tile + I.num_tiles;
        double total = 0.0;
        for (uint32_t p = 0; p < C.num_passes; ++p) {
            const BPass &P = passes[C.first_pass + p];
            const uint32_t flags = P.flags;
            uint64_t tmask = 0;
#pragma unroll
            for (int b = 0; b < TB; ++b)
                if ((tid >> b) & 1u) tmask |= P.ts.qbit[b];
            for (uint32_t t = I.first_tile; t < t_end; ++t) {
                const uint64_t base = tile_base(t, P.ts);
                const uint64_t dt = base | tmask;
                dv2 v[TR];
                if (flags & kPassInit) {
#pragma unroll
                    for (int r = 0; r < TR; ++r) v[r] = mk2(slot_offset(dt, r, P.ts) == 0 ? 1.0 : 0.0, 0.0);
                } else {
#pragma unroll
                    for (int r = 0; r < TR; ++r) v[r] = a[slot_offset(dt, r, P.ts)];
                }
                // the next gate's record is loaded while the current gate is applied
                const uint32_t ng = P.num_gates;
                if (ng) {
                    const TGate *pg = gates + P.first_gate;
                    TGate G = pg[0];
                    uint32_t i = 0;
                    do {
                        const TGate Gn = pg[i + 1 < ng ? i + 1 : i];
                        if ((base & G.ocm) == G.ocm) {
                            if (G.kind == KZ || G.kind == KP || G.kind == KRZ)
                                apply_diag(v, base, tid, G);
                            else
                                apply_pair(v, lds, tid, G);
                        }
                        G = Gn;
                    } while (++i < ng);
                }
                if (flags & kPassStore) {
#pragma unroll
                    for (int r = 0; r < TR; ++r) a[slot_offset(dt, r, P.ts)] = v[r];
                }
                if (P.num_str) {
                    // <H> = sum_k coeff_k <P_k>, <P> = sum_j conj(a_{j^x}) i^ny (-1)^{j.z} a_j
                    if (flags & kPassLds) {
                        __syncthreads();  // earlier exchange reads are done
#pragma unroll
                        for (int r = 0; r < TR; ++r) lds[tid | (r << TB)] = v[r];
                        __syncthreads();
                    }
                    double acc = 0.0;
                    for (uint32_t k = 0; k < P.num_str; ++k) {
                        const BString S = strings[P.first_str + k];
                        // popcount(i & z) = popcount(dt & z) + popcount(r & zr), zr = z on the register bits
                        uint32_t zr = 0;
#pragma unroll
                        for (int b = 0; b < TL - TB; ++b)
                            if (P.ts.qbit[TB + b] & S.z) zr |= 1u << b;
                        const bool pd = (__popcll(dt & S.z) & 1) != 0;
                        double sk = 0.0;
#pragma unroll
                        for (int r = 0; r < TR; ++r) {
                            const bool neg = pd != ((__builtin_popcount((uint32_t)r & zr) & 1) != 0);
                            double term;
                            bool flip = neg;
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
                    total += block_sum(acc, ws);
                }
            }
            __threadfence_block();
            __syncthreads();
            if (I.nchunks > 1 && p + 1 < C.num_passes) circuit_barrier(bar + ci, (p + 1) * I.nchunks, err);
        }
        if (tid == 0) {
            if (I.nchunks == 1) {
                results[ci] = total;
            } else {
                // the last chunk to finish sums the chunks' energies in chunk
                // order, so the result does not depend on which one that is
                partial[item] = total;
                __threadfence();
                if (atomicAdd(done + ci, 1u) + 1 == I.nchunks) {
                    __threadfence();
                    const uint32_t first = item - I.chunk;
                    double sum = 0.0;
                    for (uint32_t j = 0; j < I.nchunks; ++j)
                        sum += __hip_atomic_load(partial + first + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    results[ci] = sum;
                }
            }
            s_item = atomicAdd(counter, 1u);
        }
        __syncthreads();
    }
}

}  // namespace

struct kbwh_batch {
    uint32_t ncirc = 0;
    uint64_t nparams = 0;
    uint64_t ngates = 0;
    hipStream_t stream = nullptr;
    dv2 *states = nullptr;
    BCirc *d_circs = nullptr;
    BItem *d_items = nullptr;
    uint32_t nitems = 0;
    double *d_partial = nullptr;  // chunk energies, one per item
    BPass *d_passes = nullptr;
    TGate *d_tmpl = nullptr, *d_live = nullptr;
    BString *d_strings = nullptr;
    double *d_params = nullptr, *d_results = nullptr;
    uint32_t *d_ctl = nullptr;  // item counter, error flag, per-circuit barrier and done counters
    double *h_params = nullptr, *h_results = nullptr;  // pinned; h_results ends with the error flag
    std::vector<BCirc> circs;
    uint32_t npasses = 0;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t grid = 0;
};

namespace {

int32_t enqueue_run(kbwh_batch *b) {
    hipStream_t st = b->stream;
    hipError_t e = hipSuccess;
    if (b->nparams)
        e = hipMemcpyAsync(b->d_params, b->h_params, b->nparams * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->ngates) {
        k_coef<<<(uint32_t)((b->ngates + 255) / 256), 256, 0, st>>>(b->d_tmpl, b->d_live, b->ngates, b->d_params);
        e = hipGetLastError();
    }
    const uint32_t nctl = 2 + 2 * b->ncirc;
    if (e == hipSuccess) {
        k_reset<<<std::min<uint32_t>((nctl + 255) / 256, 64), 256, 0, st>>>(b->d_ctl, nctl);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_batch<<<b->grid, TT, 0, st>>>(b->states, b->d_circs, b->d_items, b->nitems, b->d_passes, b->d_live,
                                       b->d_strings, b->d_results, b->d_partial, b->d_ctl, b->ncirc);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->h_results, b->d_results, b->ncirc * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->h_results + b->ncirc, b->d_ctl + 1, 4, hipMemcpyDeviceToHost, st);
    return hip_status(e, "batch run");
}

void destroy(kbwh_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void *dev[] = {b->states, b->d_circs, b->d_items, b->d_partial, b->d_passes, b->d_tmpl, b->d_live,
                   b->d_strings, b->d_params, b->d_results, b->d_ctl};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    if (b->h_params) (void)hipHostFree(b->h_params);
    if (b->h_results) (void)hipHostFree(b->h_results);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

template <typename T>
hipError_t upload(T **dst, const std::vector<T> &src) {
    const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
    hipError_t e = hipMalloc(dst, bytes);
    if (e == hipSuccess && !src.empty()) e = hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

}  // namespace

namespace {

struct Plan {
    std::vector<BCirc> circs;
    std::vector<BPass> passes;
    std::vector<TGate> tmpl;
    std::vector<BString> strings;
    uint64_t state_total = 0;
};

// Plans every circuit's tiled passes on the host (no device needed) and
// checks the plan: every non-diagonal gate's target is a tile position, and
// every pass's tile holds TL distinct qubits of the circuit's width.
int32_t build_plan(const kbwh_circuit *circuits, uint32_t num_circuits, const kbwh_op *ops,
                   const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                   const double *coeffs, uint64_t num_params, Plan &pl) {
    std::vector<BCirc> &circs = pl.circs;
    std::vector<BPass> &passes = pl.passes;
    std::vector<TGate> &tmpl = pl.tmpl;
    std::vector<BString> &strings = pl.strings;
    circs.resize(num_circuits);
    uint64_t &state_total = pl.state_total;
    for (uint32_t c = 0; c < num_circuits; ++c) {
        const kbwh_circuit &K = circuits[c];
        const uint32_t n = K.num_qubits;
        if (n == 0 || n > kMaxBatchQubits) return set_error(KBWH_ERR_EXECUTION_FAILED, "batch circuit size out of range");
        if (K.num_ops && !ops) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null op array");
        const uint32_t n_eff = std::max<uint32_t>(n, TL);
        BCirc &B = circs[c];
        B.state_off = state_total;
        B.n_eff = n_eff;
        state_total += 1ull << n_eff;
        B.first_pass = (uint32_t)passes.size();
        // plan the circuit's passes (physical = logical qubits)
        Pass cur;
        auto close = [&]() {
            if (cur.ops.empty()) return;
            int pos_of[64];
            uint32_t q_at[TL];
            assign_positions(n_eff, cur, pos_of, q_at);
            BPass P;
            memset(&P, 0, sizeof P);
            P.ts = make_tile_spec(q_at);
            P.first_gate = (uint32_t)tmpl.size();
            P.num_gates = (uint32_t)cur.ops.size();
            for (size_t i = 0; i < cur.ops.size(); ++i) tmpl.push_back(make_tgate(cur.ops[i], pos_of, n_eff));
            passes.push_back(P);
            cur = Pass();
        };
        std::vector<int32_t> pidx;
        std::vector<PhysOp> phys;
        for (uint64_t i = K.first_op; i < K.first_op + K.num_ops; ++i) {
            const kbwh_op &op = ops[i];
            if (op.kind > KRZ || op.target >= n || (op.ctrl_mask >> n))
                return set_error(KBWH_ERR_INVALID_ARGUMENT, "batch op out of range");
            int32_t gp = -1;
            if (op.param >= 0) {
                if (!(op.kind == KRX || op.kind == KRY || op.kind == KRZ || op.kind == KP))
                    return set_error(KBWH_ERR_INVALID_ARGUMENT, "parameterised ops must be RX, RY, RZ or P");
                const uint64_t pi = (uint64_t)K.first_param + (uint64_t)op.param;
                if (pi >= num_params) return set_error(KBWH_ERR_INVALID_ARGUMENT, "param index out of range");
                gp = (int32_t)pi;
            }
            phys.push_back(PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), op.angle});
            pidx.push_back(gp);
        }
        // dependency-ordered passes (as KBWH_RUN_REORDER): batch energies
        // agree with the reference to rounding, and the order only moves
        // gates past gates on other qubits
        std::vector<uint32_t> order, ends;
        schedule_passes(phys, n_eff, TL - 3, order, ends);
        size_t k0 = 0;
        for (const uint32_t end : ends) {
            for (size_t k = k0; k < end; ++k) add_op(cur, phys[order[k]]);
            close();
            k0 = end;
        }
        const uint32_t gate_passes = (uint32_t)passes.size() - B.first_pass;
        // attach parameter indices to this circuit's gate records, in execution order
        if (gate_passes) {
            const size_t gi = passes[B.first_pass].first_gate;
            for (size_t k = 0; k < order.size(); ++k) tmpl[gi + k].pad = (uint32_t)pidx[order[k]];
        } else {
            // no gates: one empty pass prepares and stores |0...0>
            uint32_t q_at[TL];
            for (int b = 0; b < TL; ++b) q_at[b] = (uint32_t)b;
            BPass P;
            memset(&P, 0, sizeof P);
            P.ts = make_tile_spec(q_at);
            P.first_gate = (uint32_t)tmpl.size();
            passes.push_back(P);
        }
        for (uint32_t p = B.first_pass; p < (uint32_t)passes.size(); ++p) passes[p].flags = kPassStore;
        passes[B.first_pass].flags |= kPassInit;
        // the Hamiltonian: decode the strings
        std::vector<BString> hs;
        B.num_strings = K.num_strings;
        for (uint32_t k = K.first_string; k < K.first_string + K.num_strings; ++k) {
            if (!offsets || !coeffs) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null Hamiltonian arrays");
            BString S{0, 0, 0, 0, coeffs[k]};
            uint64_t seen = 0;
            for (uint32_t j = offsets[k]; j < offsets[k + 1]; ++j) {
                const uint32_t q = qubits[j];
                if (q >= n) return set_error(KBWH_ERR_INVALID_ARGUMENT, "Pauli qubit out of range");
                if (seen >> q & 1ull) return set_error(KBWH_ERR_INVALID_ARGUMENT, "qubit repeated inside one Pauli string");
                seen |= 1ull << q;
                const uint64_t bit = 1ull << q;
                if (paulis[j] == KBWH_PAULI_X) S.x |= bit;
                else if (paulis[j] == KBWH_PAULI_Y) { S.x |= bit; S.z |= bit; S.ny++; }
                else if (paulis[j] == KBWH_PAULI_Z) S.z |= bit;
                else return set_error(KBWH_ERR_INVALID_ARGUMENT, "unknown Pauli letter");
            }
            hs.push_back(S);
        }
        // strings evaluated in a pass: those whose X/Y qubits are tile qubits
        auto take = [&](BPass &P, std::vector<bool> &done) {
            uint64_t qmask = 0;
            for (int b = 0; b < TL; ++b) qmask |= P.ts.qbit[b];
            P.first_str = (uint32_t)strings.size();
            for (size_t k = 0; k < hs.size(); ++k) {
                if (done[k] || (hs[k].x & ~qmask)) continue;
                BString S = hs[k];
                S.xt = 0;
                for (int b = 0; b < TL; ++b)
                    if (S.x & P.ts.qbit[b]) S.xt |= 1u << b;
                if (S.xt) P.flags |= kPassLds;
                strings.push_back(S);
                done[k] = true;
            }
            P.num_str = (uint32_t)strings.size() - P.first_str;
        };
        std::vector<bool> done(hs.size(), false);
        take(passes.back(), done);  // the last gate pass, on the final state
        for (;;) {
            // a read-only pass for the rest: qubits 0..2, then the X/Y qubits of
            // the remaining strings in order while they fit, then filler qubits
            uint64_t qset = 7;
            bool any = false;
            for (size_t k = 0; k < hs.size(); ++k) {
                if (done[k]) continue;
                any = true;
                if (__builtin_popcountll(qset | hs[k].x) <= TL) qset |= hs[k].x;
            }
            if (!any) break;
            for (uint32_t q = 3; q < n_eff && __builtin_popcountll(qset) < TL; ++q) qset |= 1ull << q;
            uint32_t q_at[TL];
            int b = 0;
            for (uint32_t q = 0; q < n_eff; ++q)
                if (qset >> q & 1ull) q_at[b++] = q;
            BPass P;
            memset(&P, 0, sizeof P);
            P.ts = make_tile_spec(q_at);
            P.first_gate = (uint32_t)tmpl.size();
            passes.push_back(P);
            take(passes.back(), done);
            if (passes.back().num_str == 0) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: string placement");
        }
        B.num_passes = (uint32_t)passes.size() - B.first_pass;
    }
    // plan checks
    for (const BCirc &B : circs) {
        if (B.n_eff < TL || B.n_eff > kMaxBatchQubits) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: width");
        for (uint32_t p = B.first_pass; p < B.first_pass + B.num_passes; ++p) {
            const BPass &P = passes[p];
            uint64_t seen = 0;
            for (int b = 0; b < TL; ++b) {
                const uint64_t q = P.ts.qbit[b];
                if (!q || (q & (q - 1)) || (q >> B.n_eff) || (seen & q))
                    return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: tile qubits");
                seen |= q;
            }
            if (P.first_str + P.num_str > strings.size()) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: strings");
            for (uint32_t k = P.first_str; k < P.first_str + P.num_str; ++k)
                if (strings[k].xt >> TL || __builtin_popcountll(strings[k].x) != __builtin_popcount(strings[k].xt))
                    return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: string outside tile");
            for (uint32_t i = P.first_gate; i < P.first_gate + P.num_gates; ++i) {
                const TGate &g = tmpl[i];
                const bool diag = g.kind == KZ || g.kind == KP || g.kind == KRZ;
                if (!diag && g.tpos >= (uint32_t)TL) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: target outside tile");
                if (g.tpos != kOutside && g.tpos >= (uint32_t)TL) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: tile position");
                if (g.cm >> TL) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: control position");
            }
        }
    }
    return KBWH_OK;
}

}  // namespace

extern "C" {

int32_t kbwh_batch_plan(const kbwh_circuit *circuits, uint32_t num_circuits, const kbwh_op *ops,
                        const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                        const double *coeffs, uint64_t num_params, uint64_t *gates, uint64_t *passes,
                        uint64_t *amplitudes) {
    if (!gates || !passes || !amplitudes || (num_circuits && !circuits))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    Plan pl;
    int32_t rc = build_plan(circuits, num_circuits, ops, paulis, qubits, offsets, coeffs, num_params, pl);
    if (rc) return rc;
    *gates = pl.tmpl.size();
    *passes = pl.passes.size();
    *amplitudes = pl.state_total;
    return KBWH_OK;
}

int32_t kbwh_batch_new(const kbwh_circuit *circuits, uint32_t num_circuits, const kbwh_op *ops,
                       const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                       const double *coeffs, uint64_t num_params, kbwh_batch **out) {
    if (!out || (num_circuits && !circuits)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(KBWH_ERR_NO_DEVICE, "no HIP device visible");
    Plan pl;
    int32_t prc = build_plan(circuits, num_circuits, ops, paulis, qubits, offsets, coeffs, num_params, pl);
    if (prc) return prc;
    std::vector<BCirc> &circs = pl.circs;
    std::vector<BPass> &passes = pl.passes;
    std::vector<TGate> &tmpl = pl.tmpl;
    std::vector<BString> &strings = pl.strings;
    const uint64_t state_total = pl.state_total;
    // largest circuits first: the persistent workgroups balance the tail
    std::vector<uint32_t> order(num_circuits);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return circs[x].n_eff > circs[y].n_eff; });
    kbwh_batch *b = new kbwh_batch();
    b->ncirc = num_circuits;
    b->nparams = num_params;
    b->ngates = tmpl.size();
    b->circs = circs;
    b->npasses = (uint32_t)passes.size();
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // persistent grid: 4 resident workgroups per CU (launch bounds of 4 waves
    // per SIMD, 32 KiB of LDS each)
    const uint32_t resident = (uint32_t)std::max(cus, 1) * 4;
    const uint32_t max_chunks = std::min<uint32_t>(kMaxChunks, resident);
    std::vector<BItem> items;
    for (const uint32_t c : order) {
        const uint32_t ntiles = 1u << (circs[c].n_eff - TL);
        const uint32_t nch = std::max<uint32_t>(1, std::min<uint32_t>(max_chunks, ntiles / kChunkTiles));
        for (uint32_t j = 0; j < nch; ++j) {
            BItem I;
            memset(&I, 0, sizeof I);
            I.circ = c;
            I.chunk = j;
            I.nchunks = nch;
            I.first_tile = (uint32_t)((uint64_t)ntiles * j / nch);
            I.num_tiles = (uint32_t)((uint64_t)ntiles * (j + 1) / nch) - I.first_tile;
            items.push_back(I);
        }
    }
    b->nitems = (uint32_t)items.size();
    b->grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(items.size(), resident));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->states, std::max<uint64_t>(state_total, 1) * sizeof(dv2));
    if (e == hipSuccess) e = upload(&b->d_circs, circs);
    if (e == hipSuccess) e = upload(&b->d_items, items);
    if (e == hipSuccess) e = hipMalloc(&b->d_partial, std::max<size_t>(items.size(), 1) * 8);
    if (e == hipSuccess) e = upload(&b->d_passes, passes);
    if (e == hipSuccess) e = upload(&b->d_tmpl, tmpl);
    if (e == hipSuccess) e = hipMalloc(&b->d_live, std::max<size_t>(tmpl.size(), 1) * sizeof(TGate));
    if (e == hipSuccess) e = upload(&b->d_strings, strings);
    if (e == hipSuccess) e = hipMalloc(&b->d_params, std::max<uint64_t>(num_params, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_results, std::max<uint32_t>(num_circuits, 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&b->d_ctl, (2 + 2 * (size_t)num_circuits) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(&b->h_params, std::max<uint64_t>(num_params, 1) * 8, 0);
    if (e == hipSuccess) e = hipHostMalloc(&b->h_results, ((size_t)num_circuits + 1) * 8, 0);
    if (e != hipSuccess) {
        int32_t rc = hip_status(e, "kbwh_batch_new");
        destroy(b);
        return rc;
    }
    *out = b;
    return KBWH_OK;
}

int32_t kbwh_batch_run(kbwh_batch *b, const double *params, double *exp_values) {
    if (!b || !exp_values || (b->nparams && !params)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    if (b->nparams) memcpy(b->h_params, params, b->nparams * 8);
    hipError_t e = hipSuccess;
    if (b->use_graph) {
        if (!b->exec) {
            e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_status(e, "graph capture");
            int32_t rc = enqueue_run(b);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e2 != hipSuccess) return hip_status(e2, "graph capture end");
            b->graph = g;
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) return hip_status(e, "graph instantiate");
        }
        e = hipGraphLaunch(b->exec, b->stream);
        if (e != hipSuccess) return hip_status(e, "graph launch");
    } else {
        int32_t rc = enqueue_run(b);
        if (rc) return rc;
    }
    e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_status(e, "batch sync");
    uint32_t flag = 0;
    memcpy(&flag, b->h_results + b->ncirc, 4);
    if (flag) return set_error(KBWH_ERR_DEVICE, "batch: a circuit's chunks did not meet at a pass barrier in time");
    memcpy(exp_values, b->h_results, b->ncirc * 8);
    return KBWH_OK;
}

int32_t kbwh_batch_use_graph(kbwh_batch *b, int32_t enable) {
    if (!b) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    b->use_graph = enable != 0;
    return KBWH_OK;
}

int32_t kbwh_batch_state(kbwh_batch *b, uint32_t circuit, uint64_t count, double *out) {
    if (!b || !out || circuit >= b->ncirc) return set_error(KBWH_ERR_INVALID_ARGUMENT, "bad argument");
    const BCirc &C = b->circs[circuit];
    if (count > (1ull << C.n_eff)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "count out of range");
    hipError_t e = hipMemcpyAsync(out, b->states + C.state_off, count * sizeof(dv2), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return hip_status(e, "batch state");
}

int32_t kbwh_batch_stats(const kbwh_batch *b, uint64_t *gates, uint64_t *passes, uint64_t *amplitudes) {
    if (!b || !gates || !passes || !amplitudes) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *gates = b->ngates;
    *passes = b->npasses;
    uint64_t amps = 0;
    for (const BCirc &c : b->circs) amps += (1ull << c.n_eff) * c.num_passes;
    *amplitudes = amps;
    return KBWH_OK;
}

int32_t kbwh_batch_free(kbwh_batch *b) {
    destroy(b);
    return KBWH_OK;
}

}  // extern "C"
