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
#include <map>
#include <string>
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
// A structure group runs as sets when its launches cover at least this many
// tiles (one per CU): smaller groups stay in the persistent kernel.
constexpr uint64_t kMinGroupTiles = 256;

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
    // the same trip count in every lane (no per-lane loop exit in this unit)
    for (uint32_t i0 = blockIdx.x * 256; i0 < count; i0 += gridDim.x * 256) {
        const uint32_t i = i0 + threadIdx.x;
        if (i < count) ctl[i] = 0;
    }
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

// ---- groups of circuits that share one structure ----------------------
//
// Circuits whose gates, controls, parameter slots and Pauli strings are the
// same (only parameters and string weights differ) run as one state of
// sets: S sets of 2^n_eff amplitudes one after another, one launch per pass
// over all of them (per-set gate records: tile t takes the records of set
// t >> (n_eff - TL)), then read sweeps for the strings with per-set weights.
// Every launch covers S x 2^(n_eff - TL) tiles, so large groups fill the
// device the way one large state does, with no barriers between workgroups.

// per-set records: out[dst[r] + s stride[r]] = template r with the
// coefficients of set s's parameters (pad: parameter slot relative to the
// circuit's first parameter, or -1)
__global__ __launch_bounds__(256) void k_expand_group(const TGate *__restrict__ tmpl, const uint32_t *__restrict__ dst,
                                                      const uint32_t *__restrict__ stride, uint32_t nrec,
                                                      uint32_t nsets, const uint32_t *__restrict__ pbase,
                                                      const double *__restrict__ params, TGate *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)nrec * nsets) return;
    const uint32_t s = (uint32_t)(i / nrec), r = (uint32_t)(i % nrec);
    TGate g = tmpl[r];
    const int32_t p = (int32_t)g.pad;
    if (p >= 0) set_coefficients_dev(g, params[(uint64_t)pbase[s] + (uint32_t)p]);
    out[dst[r] + (uint64_t)s * stride[r]] = g;
}

// one pass of a group with the generic gate loop (passes without a compiled
// kernel); INIT: the tile starts as |0...0> of its set instead of a load
template <bool INIT>
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void k_gpass(
    dv2 *__restrict__ a, TileSpec ts, const TGate *__restrict__ gates, uint32_t ng, uint32_t ntiles,
    uint32_t set_shift, uint64_t set_mask) {
    __shared__ dv2 lds[TT * kChunk];
    const uint32_t tid = threadIdx.x, tile = blockIdx.x;
    if (tile >= ntiles) return;
    gates += (size_t)(tile >> set_shift) * ng;
    const uint64_t base = tile_base(tile, ts);
    const uint64_t dt = thread_offset(base, tid, ts);
    dv2 v[TR];
    if constexpr (INIT)
        tile_zero_state(v, dt, ts, set_mask);
    else
        tile_load(v, a, dt, ts);
    if (ng) {
        const TGate *gp = gates, *const ge = gates + ng;
        do {
            const TGate G = *gp;
            apply_gate(v, lds, base, tid, G);
        } while (++gp != ge);
    }
    tile_store(v, a, dt, ts);
}

// one read sweep: the strings [first, first + ns) of the group's table, whose
// X/Y qubits are tile qubits of ts, weighted per set (coef[set * nstr + k]);
// partial[tile] = the tile's sum (arithmetic as in k_batch)
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(KBWH_TILE_WAVES, KBWH_TILE_WAVES))) void k_gsweep(
    const dv2 *__restrict__ a, TileSpec ts, const BString *__restrict__ strings, uint32_t first, uint32_t ns,
    uint32_t has_x, uint32_t ntiles, uint32_t set_shift, const double *__restrict__ coef, uint32_t nstr,
    double *__restrict__ partial) {
    __shared__ dv2 lds[TT * TR];
    __shared__ double ws[TT / 64];
    const uint32_t tid = threadIdx.x, tile = blockIdx.x;
    if (tile >= ntiles) return;
    const double *cw = coef + (size_t)(tile >> set_shift) * nstr + first;
    const uint64_t dt = thread_offset(tile_base(tile, ts), tid, ts);
    dv2 v[TR];
    tile_load(v, a, dt, ts);
    if (has_x) {
#pragma unroll
        for (int r = 0; r < TR; ++r) lds[tid | (r << TB)] = v[r];
        __syncthreads();
    }
    double acc = 0.0;
    for (uint32_t k = 0; k < ns; ++k) {
        const BString S = strings[first + k];
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
                const uint32_t q = S.ny & 3u;
                term = (q & 1u) ? wi : wr;
                flip = flip != (((q + 1u) & 2u) != 0);
            }
            sk += flip ? -term : term;
        }
        acc += cw[k] * sk;
    }
    const double total = block_sum(acc, ws);
    if (tid == 0) partial[tile] = total;
}

// one workgroup per set: the set's tiles over every sweep, in a fixed order
// (each thread a fixed stride of tiles, then the block sum), into the
// circuit's result
__global__ __launch_bounds__(256) void k_gsum(const double *__restrict__ partial, uint32_t nsweeps, uint32_t ntiles,
                                              uint32_t tiles_per_set, const uint32_t *__restrict__ circ,
                                              double *__restrict__ results) {
    __shared__ double ws[256 / 64];
    const uint32_t s = blockIdx.x;
    double acc = 0.0;
    for (uint32_t g = 0; g < nsweeps; ++g) {
        const double *p = partial + (size_t)g * ntiles + (size_t)s * tiles_per_set;
        for (uint32_t t0 = 0; t0 < tiles_per_set; t0 += 256) {
            const uint32_t t = t0 + threadIdx.x;
            if (t < tiles_per_set) acc += p[t];
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) results[circ[s]] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

struct GPass {
    TileSpec ts;
    uint32_t first, count;  // template records [first, first + count)
    uint32_t flags;         // kJitInit on the first pass
    uint32_t pad;
};

struct GSweep {
    TileSpec ts;
    uint32_t first, count, has_x, pad;
};

struct SetGroup {
    uint32_t n_eff = 0, nsets = 0, cap = TL - 3;
    bool compiled = false;
    double tuned_ms = 0.0;  // the chosen configuration's time in the last tuning (0: not tuned)
    uint64_t state_off = 0;
    std::vector<uint32_t> circ;  // circuit of each set
    std::vector<GPass> passes;
    std::vector<TGate> tmpl;
    std::vector<GSweep> sweeps;
    uint32_t nstr = 0;
    TGate *d_tmpl = nullptr, *d_recs = nullptr;
    uint32_t *d_meta = nullptr;  // dst[nrec], stride[nrec], pbase[nsets], circ[nsets]
    BString *d_str = nullptr;
    double *d_coef = nullptr, *d_partial = nullptr;
    std::vector<void *> fns;  // compiled kernel of each pass (null: k_gpass)
};

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
    std::vector<SetGroup> groups;
    int device = 0;
    // the inputs, kept for re-planning groups (kbwh_batch_configure / _autotune)
    std::vector<kbwh_circuit> in_circuits;
    std::vector<kbwh_op> in_ops;
    std::vector<uint8_t> in_paulis;
    std::vector<uint32_t> in_qubits, in_offsets;
    std::vector<double> in_coeffs;
};

namespace {

hipError_t enqueue_group(const SetGroup &G, dv2 *states, const double *params, double *results, hipStream_t st);

int32_t enqueue_run(kbwh_batch *b) {
    hipStream_t st = b->stream;
    hipError_t e = hipSuccess;
    if (b->nparams)
        e = hipMemcpyAsync(b->d_params, b->h_params, b->nparams * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && b->ngates && b->nitems) {
        k_coef<<<(uint32_t)((b->ngates + 255) / 256), 256, 0, st>>>(b->d_tmpl, b->d_live, b->ngates, b->d_params);
        e = hipGetLastError();
    }
    const uint32_t nctl = 2 + 2 * b->ncirc;
    if (e == hipSuccess) {
        k_reset<<<std::min<uint32_t>((nctl + 255) / 256, 64), 256, 0, st>>>(b->d_ctl, nctl);
        e = hipGetLastError();
    }
    if (e == hipSuccess && b->nitems) {
        k_batch<<<b->grid, TT, 0, st>>>(b->states, b->d_circs, b->d_items, b->nitems, b->d_passes, b->d_live,
                                       b->d_strings, b->d_results, b->d_partial, b->d_ctl, b->ncirc);
        e = hipGetLastError();
    }
    for (size_t g = 0; g < b->groups.size() && e == hipSuccess; ++g)
        e = enqueue_group(b->groups[g], b->states, b->d_params, b->d_results, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->h_results, b->d_results, b->ncirc * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->h_results + b->ncirc, b->d_ctl + 1, 4, hipMemcpyDeviceToHost, st);
    return hip_status(e, "batch run");
}

void free_group(SetGroup &G);

void destroy(kbwh_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (SetGroup &G : b->groups) free_group(G);
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
    hipError_t e = dev_alloc(dst, bytes);
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

// The structure of a circuit: width, ops (kind, target, controls, parameter
// slot, and the angle of an unparameterised op) and the letters and qubits of
// its Pauli strings.  Circuits with equal keys differ only in parameters and
// string weights.
std::string circuit_key(const kbwh_circuit &K, const kbwh_op *ops, const uint8_t *paulis, const uint32_t *qubits,
                        const uint32_t *offsets) {
    std::string k;
    auto put = [&](const void *p, size_t n) { k.append((const char *)p, n); };
    put(&K.num_qubits, sizeof K.num_qubits);
    put(&K.num_ops, sizeof K.num_ops);
    put(&K.num_strings, sizeof K.num_strings);
    for (uint64_t i = K.first_op; i < K.first_op + K.num_ops; ++i) {
        const kbwh_op &op = ops[i];
        put(&op.kind, sizeof op.kind);
        put(&op.target, sizeof op.target);
        put(&op.ctrl_mask, sizeof op.ctrl_mask);
        put(&op.param, sizeof op.param);
        // a fixed angle counts for P, RX, RY and RZ only (X, Y, Z and H
        // ignore it), by value (-0.0 and 0.0 are one key)
        if (op.param < 0 && (op.kind == KP || op.kind == KRX || op.kind == KRY || op.kind == KRZ)) {
            const double a = op.angle == 0.0 ? 0.0 : op.angle;
            put(&a, sizeof a);
        }
    }
    for (uint32_t j = K.first_string; j < K.first_string + K.num_strings; ++j) {
        const uint32_t len = offsets[j + 1] - offsets[j];
        put(&len, sizeof len);
        for (uint32_t t = offsets[j]; t < offsets[j + 1]; ++t) {
            put(&paulis[t], 1);
            put(&qubits[t], sizeof qubits[t]);
        }
    }
    return k;
}

// Plans a group's passes (dependency order, at most `cap` non-diagonal
// targets above qubit 2 per pass) and read sweeps from its first circuit.
// Inputs were checked by build_plan.
int32_t plan_group(SetGroup &G, uint32_t cap, const kbwh_circuit *circuits, const kbwh_op *ops, const uint8_t *paulis,
                   const uint32_t *qubits, const uint32_t *offsets, const double *coeffs) {
    const kbwh_circuit &K = circuits[G.circ[0]];
    const uint32_t n = G.n_eff;
    G.cap = cap;
    G.passes.clear();
    G.tmpl.clear();
    G.sweeps.clear();
    std::vector<PhysOp> phys;
    std::vector<int32_t> slot;
    for (uint64_t i = K.first_op; i < K.first_op + K.num_ops; ++i) {
        const kbwh_op &op = ops[i];
        phys.push_back(PhysOp{(int)op.kind, op.target, op.ctrl_mask & ~(1ull << op.target), op.angle});
        slot.push_back(op.param);
    }
    std::vector<uint32_t> order, ends;
    schedule_passes(phys, n, cap, order, ends);
    if (order.size() != phys.size()) return set_error(KBWH_ERR_DEVICE, "group schedule incomplete");
    size_t k0 = 0;
    for (const uint32_t end : ends) {
        Pass cur;
        for (size_t k = k0; k < end; ++k) add_op(cur, phys[order[k]]);
        int pos_of[64];
        uint32_t q_at[TL];
        assign_positions(n, cur, pos_of, q_at);
        GPass P;
        memset(&P, 0, sizeof P);
        P.ts = make_tile_spec(q_at);
        P.first = (uint32_t)G.tmpl.size();
        P.count = (uint32_t)(end - k0);
        for (size_t k = k0; k < end; ++k) {
            TGate g = make_tgate(phys[order[k]], pos_of, n);
            g.pad = (uint32_t)slot[order[k]];
            G.tmpl.push_back(g);
        }
        G.passes.push_back(P);
        k0 = end;
    }
    if (G.passes.empty()) {
        // no gates: one pass makes and stores |0...0>
        uint32_t q_at[TL];
        for (int b = 0; b < TL; ++b) q_at[b] = (uint32_t)b;
        GPass P;
        memset(&P, 0, sizeof P);
        P.ts = make_tile_spec(q_at);
        G.passes.push_back(P);
    }
    G.passes[0].flags = kJitInit;
    // read sweeps: qubits 0..2, then the X/Y qubits of the remaining strings
    // in order while they fit, then the lowest free qubits
    std::vector<uint64_t> xs(K.num_strings);
    for (uint32_t j = 0; j < K.num_strings; ++j) {
        uint64_t x = 0;
        for (uint32_t t = offsets[K.first_string + j]; t < offsets[K.first_string + j + 1]; ++t)
            if (paulis[t] == KBWH_PAULI_X || paulis[t] == KBWH_PAULI_Y) x |= 1ull << qubits[t];
        xs[j] = x;
    }
    std::vector<bool> done(K.num_strings, false);
    std::vector<uint32_t> sorder;  // string of each table entry, in sweep order
    for (;;) {
        uint64_t qset = 7;
        bool any = false;
        for (uint32_t j = 0; j < K.num_strings; ++j) {
            if (done[j]) continue;
            any = true;
            if (__builtin_popcountll(qset | xs[j]) <= TL) qset |= xs[j];
        }
        if (!any) break;
        for (uint32_t q = 3; q < n && __builtin_popcountll(qset) < TL; ++q) qset |= 1ull << q;
        uint32_t q_at[TL];
        int b = 0;
        for (uint32_t q = 0; q < n; ++q)
            if (qset >> q & 1ull) q_at[b++] = q;
        GSweep W;
        memset(&W, 0, sizeof W);
        W.ts = make_tile_spec(q_at);
        W.first = (uint32_t)sorder.size();
        for (uint32_t j = 0; j < K.num_strings; ++j) {
            if (done[j] || (xs[j] & ~qset)) continue;
            sorder.push_back(j);
            done[j] = true;
            if (xs[j]) W.has_x = 1;
        }
        W.count = (uint32_t)sorder.size() - W.first;
        if (W.count == 0) return set_error(KBWH_ERR_INVALID_ARGUMENT, "plan: string placement");
        G.sweeps.push_back(W);
    }
    G.nstr = (uint32_t)sorder.size();
    // device tables: records, metadata, strings, per-set weights
    const uint32_t S = G.nsets, nrec = (uint32_t)G.tmpl.size();
    std::vector<uint32_t> meta(2 * (size_t)nrec + 2 * (size_t)S);
    for (const GPass &P : G.passes)
        for (uint32_t r = P.first; r < P.first + P.count; ++r) {
            meta[r] = S * P.first + (r - P.first);
            meta[nrec + r] = P.count;
        }
    std::vector<BString> strs;
    for (uint32_t e = 0; e < G.nstr; ++e) {
        const uint32_t j = sorder[e];
        BString B{0, 0, 0, 0, 0.0};
        for (uint32_t t = offsets[K.first_string + j]; t < offsets[K.first_string + j + 1]; ++t) {
            const uint64_t bit = 1ull << qubits[t];
            if (paulis[t] == KBWH_PAULI_X) B.x |= bit;
            else if (paulis[t] == KBWH_PAULI_Y) { B.x |= bit; B.z |= bit; B.ny++; }
            else B.z |= bit;
        }
        strs.push_back(B);
    }
    for (const GSweep &W : G.sweeps)
        for (uint32_t e = W.first; e < W.first + W.count; ++e)
            for (int b = 0; b < TL; ++b)
                if (strs[e].x & W.ts.qbit[b]) strs[e].xt |= 1u << b;
    std::vector<double> coef((size_t)S * std::max<uint32_t>(G.nstr, 1), 0.0);
    for (uint32_t s = 0; s < S; ++s) {
        const kbwh_circuit &C = circuits[G.circ[s]];
        meta[2 * (size_t)nrec + s] = C.first_param;
        meta[2 * (size_t)nrec + S + s] = G.circ[s];
        for (uint32_t e = 0; e < G.nstr; ++e) coef[(size_t)s * G.nstr + e] = coeffs[C.first_string + sorder[e]];
    }
    const uint64_t ntiles = (uint64_t)S << (n - TL);
    void *bufs[] = {G.d_tmpl, G.d_recs, G.d_meta, G.d_str, G.d_coef, G.d_partial};
    for (void *q : bufs)
        if (q) (void)hipFree(q);
    G.d_tmpl = G.d_recs = nullptr;
    G.d_meta = nullptr;
    G.d_str = nullptr;
    G.d_coef = G.d_partial = nullptr;
    hipError_t e = dev_alloc(&G.d_tmpl, std::max<size_t>(nrec, 1) * sizeof(TGate));
    if (e == hipSuccess && nrec) e = hipMemcpy(G.d_tmpl, G.tmpl.data(), nrec * sizeof(TGate), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dev_alloc(&G.d_recs, std::max<size_t>((size_t)nrec * S, 1) * sizeof(TGate));
    if (e == hipSuccess) e = dev_alloc(&G.d_meta, meta.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(G.d_meta, meta.data(), meta.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dev_alloc(&G.d_str, std::max<size_t>(strs.size(), 1) * sizeof(BString));
    if (e == hipSuccess && !strs.empty())
        e = hipMemcpy(G.d_str, strs.data(), strs.size() * sizeof(BString), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dev_alloc(&G.d_coef, coef.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(G.d_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dev_alloc(&G.d_partial, std::max<uint64_t>(ntiles * G.sweeps.size(), 1) * sizeof(double));
    G.fns.assign(G.passes.size(), nullptr);
    return hip_status(e, "batch group tables");
}

// Compiles (or drops) the group's pass-specialised kernels.
int32_t group_kernels(SetGroup &G, int device, bool compiled) {
    G.compiled = compiled;
    G.fns.assign(G.passes.size(), nullptr);
    if (!compiled) return KBWH_OK;
    std::vector<JitPass> lst;
    for (const GPass &P : G.passes) lst.push_back(JitPass{G.tmpl.data() + P.first, P.count, P.flags});
    return jit_prepare(device, lst, G.fns);
}

void free_group(SetGroup &G) {
    void *bufs[] = {G.d_tmpl, G.d_recs, G.d_meta, G.d_str, G.d_coef, G.d_partial};
    for (void *q : bufs)
        if (q) (void)hipFree(q);
    G.d_tmpl = G.d_recs = nullptr;
    G.d_meta = nullptr;
    G.d_str = nullptr;
    G.d_coef = G.d_partial = nullptr;
}

// A group's run: per-set records, its passes, its sweeps, per-set energies.
hipError_t enqueue_group(const SetGroup &G, dv2 *states, const double *params, double *results, hipStream_t st) {
    const uint32_t S = G.nsets, nrec = (uint32_t)G.tmpl.size(), n = G.n_eff;
    const uint32_t shift = n - TL, ntiles = S << shift;
    const uint64_t set_mask = (1ull << n) - 1;
    dv2 *a = states + G.state_off;
    hipError_t e = hipSuccess;
    if (nrec) {
        const uint64_t total = (uint64_t)nrec * S;
        k_expand_group<<<(uint32_t)((total + 255) / 256), 256, 0, st>>>(G.d_tmpl, G.d_meta, G.d_meta + nrec, nrec, S,
                                                                       G.d_meta + 2 * (size_t)nrec, params, G.d_recs);
        e = hipGetLastError();
    }
    for (size_t p = 0; p < G.passes.size() && e == hipSuccess; ++p) {
        const GPass &P = G.passes[p];
        const TGate *recs = G.d_recs + (size_t)S * P.first;
        if (G.fns[p]) {
            e = jit_launch(G.fns[p], (double2 *)a, &P.ts, recs, ntiles, st, shift, set_mask);
        } else {
            if (P.flags & kJitInit)
                k_gpass<true><<<ntiles, TT, 0, st>>>(a, P.ts, recs, P.count, ntiles, shift, set_mask);
            else
                k_gpass<false><<<ntiles, TT, 0, st>>>(a, P.ts, recs, P.count, ntiles, shift, set_mask);
            e = hipGetLastError();
        }
    }
    for (size_t g = 0; g < G.sweeps.size() && e == hipSuccess; ++g) {
        const GSweep &W = G.sweeps[g];
        k_gsweep<<<ntiles, TT, 0, st>>>(a, W.ts, G.d_str, W.first, W.count, W.has_x, ntiles, shift, G.d_coef, G.nstr,
                                        G.d_partial + (size_t)g * ntiles);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_gsum<<<S, 256, 0, st>>>(G.d_partial, (uint32_t)G.sweeps.size(), ntiles, 1u << shift,
                                  G.d_meta + 2 * (size_t)nrec + S, results);
        e = hipGetLastError();
    }
    return e;
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
    // circuits of one structure form groups of sets (large enough to fill
    // the device); the others run in the persistent kernel
    std::map<std::string, std::vector<uint32_t>> by_key;
    for (uint32_t c = 0; c < num_circuits; ++c)
        by_key[circuit_key(circuits[c], ops, paulis, qubits, offsets)].push_back(c);
    std::vector<SetGroup> groups;
    std::vector<uint8_t> grouped(num_circuits, 0);
    for (const auto &kv : by_key) {
        const std::vector<uint32_t> &list = kv.second;
        const uint32_t shift = circs[list[0]].n_eff - TL;
        // a launch covers at most 2^24 - 1 tiles of 256 threads
        const size_t max_sets = std::max<size_t>(1, ((1u << 24) - 1) >> shift);
        for (size_t i0 = 0; i0 < list.size(); i0 += max_sets) {
            const size_t cnt = std::min(max_sets, list.size() - i0);
            if (cnt < 2 || ((uint64_t)cnt << shift) < kMinGroupTiles) continue;
            SetGroup G;
            G.n_eff = circs[list[0]].n_eff;
            G.nsets = (uint32_t)cnt;
            G.circ.assign(list.begin() + i0, list.begin() + i0 + cnt);
            for (const uint32_t c : G.circ) grouped[c] = 1;
            groups.push_back(std::move(G));
        }
    }
    // state layout: the persistent kernel's circuits, then each group's sets
    {
        uint64_t off = 0;
        for (uint32_t c = 0; c < num_circuits; ++c)
            if (!grouped[c]) {
                circs[c].state_off = off;
                off += 1ull << circs[c].n_eff;
            }
        for (SetGroup &G : groups) {
            G.state_off = off;
            for (const uint32_t c : G.circ) {
                circs[c].state_off = off;
                off += 1ull << G.n_eff;
            }
        }
    }
    // largest circuits first: the persistent workgroups balance the tail
    std::vector<uint32_t> order;
    for (uint32_t c = 0; c < num_circuits; ++c)
        if (!grouped[c]) order.push_back(c);
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
    if (e == hipSuccess) e = dev_alloc(&b->states, std::max<uint64_t>(state_total, 1) * sizeof(dv2));
    if (e == hipSuccess) e = upload(&b->d_circs, circs);
    if (e == hipSuccess) e = upload(&b->d_items, items);
    if (e == hipSuccess) e = dev_alloc(&b->d_partial, std::max<size_t>(items.size(), 1) * 8);
    if (e == hipSuccess) e = upload(&b->d_passes, passes);
    if (e == hipSuccess) e = upload(&b->d_tmpl, tmpl);
    if (e == hipSuccess) e = dev_alloc(&b->d_live, std::max<size_t>(tmpl.size(), 1) * sizeof(TGate));
    if (e == hipSuccess) e = upload(&b->d_strings, strings);
    if (e == hipSuccess) e = dev_alloc(&b->d_params, std::max<uint64_t>(num_params, 1) * 8);
    if (e == hipSuccess) e = dev_alloc(&b->d_results, std::max<uint32_t>(num_circuits, 1) * 8);
    if (e == hipSuccess) e = dev_alloc(&b->d_ctl, (2 + 2 * (size_t)num_circuits) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc(&b->h_params, std::max<uint64_t>(num_params, 1) * 8, 0);
    if (e == hipSuccess) e = hipHostMalloc(&b->h_results, ((size_t)num_circuits + 1) * 8, 0);
    if (e != hipSuccess) {
        int32_t rc = hip_status(e, "kbwh_batch_new");
        destroy(b);
        return rc;
    }
    b->device = dev;
    b->groups = std::move(groups);
    if (!b->groups.empty()) {
        uint64_t nops = 0;
        uint32_t nstr = 0;
        for (uint32_t c = 0; c < num_circuits; ++c) {
            nops = std::max<uint64_t>(nops, circuits[c].first_op + circuits[c].num_ops);
            nstr = std::max<uint32_t>(nstr, circuits[c].first_string + circuits[c].num_strings);
        }
        const uint32_t nlet = nstr ? offsets[nstr] : 0;
        b->in_circuits.assign(circuits, circuits + num_circuits);
        b->in_ops.assign(ops, ops + nops);
        if (nlet) {
            b->in_paulis.assign(paulis, paulis + nlet);
            b->in_qubits.assign(qubits, qubits + nlet);
        }
        b->in_offsets.assign(offsets, offsets + nstr + 1);
        if (nstr) b->in_coeffs.assign(coeffs, coeffs + nstr);
    }
    for (SetGroup &G : b->groups) {
        int32_t rc = plan_group(G, TL - 3, circuits, ops, paulis, qubits, offsets, coeffs);
        if (rc) {
            destroy(b);
            return rc;
        }
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
    std::vector<uint8_t> grouped(b->circs.size(), 0);
    uint64_t amps = 0;
    uint64_t np = 0;
    for (const SetGroup &G : b->groups) {
        for (const uint32_t c : G.circ) grouped[c] = 1;
        const uint64_t per = G.passes.size() + G.sweeps.size();
        amps += ((uint64_t)G.nsets << G.n_eff) * per;
        np += per;
    }
    for (size_t c = 0; c < b->circs.size(); ++c)
        if (!grouped[c]) {
            amps += (1ull << b->circs[c].n_eff) * b->circs[c].num_passes;
            np += b->circs[c].num_passes;
        }
    *passes = np;
    *amplitudes = amps;
    return KBWH_OK;
}

int32_t kbwh_batch_free(kbwh_batch *b) {
    destroy(b);
    return KBWH_OK;
}

int32_t kbwh_batch_groups(const kbwh_batch *b, kbwh_batch_group_info *info, uint32_t capacity,
                          uint32_t *num_groups) {
    if (!b || !num_groups || (capacity && !info)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    *num_groups = (uint32_t)b->groups.size();
    for (uint32_t g = 0; g < std::min<uint32_t>(capacity, (uint32_t)b->groups.size()); ++g) {
        const SetGroup &G = b->groups[g];
        kbwh_batch_group_info &I = info[g];
        I.num_qubits = G.n_eff;
        I.num_sets = G.nsets;
        I.tile_targets = G.cap;
        I.compiled = 0;
        for (void *f : G.fns) I.compiled += f != nullptr;
        I.passes = (uint32_t)G.passes.size();
        I.sweeps = (uint32_t)G.sweeps.size();
        I.tuned_ms = G.tuned_ms;
    }
    return KBWH_OK;
}

}  // extern "C"

namespace {

// The graph holds the groups' tables and kernels: a new configuration
// drops it, and the next run captures again.
void drop_graph(kbwh_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int32_t configure_group(kbwh_batch *b, SetGroup &G, uint32_t cap, bool compiled) {
    if (G.cap != cap || G.passes.empty()) {
        int32_t rc = plan_group(G, cap, b->in_circuits.data(), b->in_ops.data(), b->in_paulis.data(),
                                b->in_qubits.data(), b->in_offsets.data(), b->in_coeffs.data());
        if (rc) return rc;
    }
    return group_kernels(G, b->device, compiled);
}

// one group's run, timed on the batch stream (after one untimed run)
int32_t time_group(kbwh_batch *b, const SetGroup &G, double *ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = enqueue_group(G, b->states, b->d_params, b->d_results, b->stream);
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    if (e == hipSuccess) e = enqueue_group(G, b->states, b->d_params, b->d_results, b->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    *ms = t;
    return hip_status(e, "batch group timing");
}

}  // namespace

extern "C" {

int32_t kbwh_batch_configure(kbwh_batch *b, uint32_t tile_targets, int32_t compiled) {
    if (!b) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null handle");
    if (tile_targets < 1 || tile_targets > (uint32_t)(TL - 3))
        return set_error(KBWH_ERR_INVALID_ARGUMENT, "tile targets must be 1..8");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_status(e, "batch sync");
    drop_graph(b);
    for (SetGroup &G : b->groups) {
        int32_t rc = configure_group(b, G, tile_targets, compiled != 0);
        if (rc) return rc;
        G.tuned_ms = 0.0;
    }
    return KBWH_OK;
}

int32_t kbwh_batch_autotune(kbwh_batch *b, const double *params, double *ms_out, uint32_t capacity) {
    if (!b || (b->nparams && !params) || (capacity && !ms_out)) return set_error(KBWH_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_status(e, "batch sync");
    drop_graph(b);
    if (b->nparams) {
        e = hipMemcpy(b->d_params, params, b->nparams * 8, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_status(e, "batch tune parameters");
    }
    static const uint32_t caps[] = {TL - 3, 6, 4};
    constexpr uint32_t kCand = 4;  // generic kernel at each cap, then compiled at the best cap
    for (size_t g = 0; g < b->groups.size(); ++g) {
        SetGroup &G = b->groups[g];
        double best = 0.0;
        uint32_t best_cap = TL - 3;
        bool best_comp = false;
        double t[kCand] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i) {
            int32_t rc = configure_group(b, G, caps[i], false);
            if (!rc) rc = time_group(b, G, &t[i]);
            if (rc) return rc;
            if (i == 0 || t[i] < best) {
                best = t[i];
                best_cap = caps[i];
            }
        }
        int32_t rc = configure_group(b, G, best_cap, true);
        bool any = false;
        for (void *f : G.fns) any = any || f != nullptr;
        if (!rc && any) rc = time_group(b, G, &t[3]);
        if (rc) return rc;
        if (any && t[3] < best) {
            best = t[3];
            best_comp = true;
        }
        rc = configure_group(b, G, best_cap, best_comp);
        if (rc) return rc;
        G.tuned_ms = best;
        for (uint32_t i = 0; i < kCand && g * kCand + i < capacity; ++i) ms_out[g * kCand + i] = t[i];
    }
    return KBWH_OK;
}

}  // extern "C"
