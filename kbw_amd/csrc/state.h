// The state handle and the host helpers shared by the engine's translation units.
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/kbwh.h"
#include "internal.h"

constexpr uint32_t kResultSlots = 4096;

// One physical operation of a Block state's gate queue, recorded when the gate
// is given (the map and LRU decisions are made then, as in block.rs:123-156):
// a gate on physical qubits, or (swap_local != 0xFFFFFFFF) the exchange of
// physical qubits swap_local and swap_global that brings a global target down.
struct kbwh_phys_event {
    int32_t kind;
    uint32_t target;
    uint64_t ctrl;  // physical control mask without the target bit
    double angle;
    uint32_t swap_local, swap_global;
};

// Device buffers of a batched adjoint gradient (kbwh_adjoint_batch), grown
// as needed and freed with the handle.
struct kbwh_set_buffers {
    void *vecs = nullptr;  // phi and lambda of every set
    size_t vecs_cap = 0;
    void *gates = nullptr;  // expanded gate records of every set
    size_t gates_cap = 0;
    void *tmpl = nullptr;  // templates, parameter indices, destinations, strides
    size_t tmpl_cap = 0;
    void *params = nullptr;  // the sets' parameters, then energies and gradient terms
    size_t params_cap = 0;
    void *partial = nullptr;  // reduction partials
    size_t partial_cap = 0;
};

struct kbwh_state {
    uint32_t n = 0;
    bool block = false;
    uint32_t block_size_arg = 0;
    uint32_t nlocal = 0;  // Block: local qubits (physical bits [0, nlocal))
    uint64_t size = 0;
    int device = 0;
    double2 *amp = nullptr;
    hipStream_t stream = nullptr;
    // reductions: per-block partials, device results, pinned host copy
    double *partial = nullptr;
    double *result = nullptr;
    double *host_result = nullptr;
    // growable device scratch (sampling levels, dump offsets, outputs)
    double *scratch = nullptr;
    size_t scratch_cap = 0;
    // Block mapping (block.rs:94-98): logical <-> physical bits and LRU stamps
    std::vector<uint32_t> l2p, p2l;
    std::vector<uint64_t> last_access;
    uint64_t counter = 0;
    uint64_t swaps = 0;
    // tiled passes: device copy of the pass gate table
    void *tile_gates = nullptr;
    size_t tile_gates_cap = 0;
    uint64_t last_passes = 0;    // HBM passes of the last kbwh_run
    // run configuration (kbwh_set_run_config / kbwh_autotune)
    uint32_t fusion_width = 5;   // qubits per fused block
    uint32_t tile_targets = 64;  // most non-diagonal targets above qubit 2 per tiled pass (64: the tile's free positions)
    bool compiled_passes = false;  // kbwh_set_compiled: the adjoint's passes on kernels compiled per pass
    int32_t plan_trials = -1;       // kbwh_set_plan_search: randomised plans tried (-1: 0 reorder, 256 compiled)
    uint32_t auto_mode = KBWH_RUN_TILED;
    // live gates queued with KBWH_FLAG_QUEUE, applied before the state is read
    bool use_queue = false;
    std::vector<kbwh_op> queue;               // flat states
    std::vector<kbwh_phys_event> block_queue;  // Block states
    uint64_t last_gates = 0;     // gates of the last kbwh_run
    // adjoint differentiation: two extra state vectors, allocated on first use
    double2 *adj_phi = nullptr;
    double2 *adj_lam = nullptr;
    double *adj_partial = nullptr;
    size_t adj_partial_cap = 0;
    kbwh_set_buffers sets;  // batched adjoint gradients
    // the last adjoint gradient's HBM passes: forward passes over phi, sweeps
    // building lambda = H phi, reverse passes over phi and lambda
    uint64_t adj_forward = 0, adj_lambda = 0, adj_reverse = 0;
};

namespace kbwh {

int32_t set_error(int32_t code, const std::string &msg);
// Every device allocation of the library goes through dev_alloc: hipMalloc,
// unless an injected failure (kbwh_inject_alloc_failure) is due, in which
// case it returns hipErrorOutOfMemory without allocating.
hipError_t dev_alloc_raw(void **p, size_t bytes);
void inject_alloc_failure(int64_t after);
template <typename T>
inline hipError_t dev_alloc(T **p, size_t bytes) {
    return dev_alloc_raw((void **)p, bytes);
}
int32_t hip_status(hipError_t e, const char *what);
int32_t ensure_scratch(kbwh_state *s, size_t doubles);
int32_t local_qubit(kbwh_state *s, uint32_t logical, uint32_t *out);
constexpr uint32_t kNoSwap = 0xFFFFFFFFu;
void map_qubit(kbwh_state *s, uint32_t logical, uint32_t *out, uint32_t *swap_local, uint32_t *swap_global);
uint64_t physical_mask(const kbwh_state *s, uint64_t logical_mask);
int32_t apply_gate(kbwh_state *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask);
int32_t reduce_to_host(kbwh_state *s, double *out, uint32_t count);
double uniform_scale(double total);
int32_t run_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params);
// One pass of the tiled kernel k_tile over a state: spec / gates are a
// tile::TileSpec and a device array of tile::TGate of the 11-qubit tile (the
// records are plain data of these sizes).  launch_tile_exact (tiled_pass.hip)
// stores the tile through the TileSpec `store` when it is not null (Block
// swaps folded into the pass); launch_tile_contracted (tiled_fma.hip) is the
// FP-contracted copy for dependency-ordered runs of flat states.
constexpr size_t kTileSpecBytes = 176, kTGateBytes = 144;
hipError_t launch_tile_exact(double2 *amp, const void *spec, const void *store, const void *gates,
                             uint32_t num_gates, uint32_t ntiles, hipStream_t stream);
// set_shift < 32: tile t uses the records gates + (t >> set_shift) * num_gates
// (parameter sets laid out one after another, see run_sets_forward).
hipError_t launch_tile_contracted(double2 *amp, const void *spec, const void *gates, uint32_t num_gates,
                                  uint32_t ntiles, hipStream_t stream, uint32_t set_shift = 32);
// reorder: dependency-ordered passes (KBWH_RUN_REORDER; flat states only);
// compiled: their pass-specialised kernels (KBWH_RUN_COMPILED)
int32_t run_tiled_on(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops,
                     const double *params, bool reorder = false, bool compiled = false);
// Pass-specialised kernels (tile_jit.hip): compiles and loads on `device` the
// kernel of each pass (cached by the pass's structure and flags); fns[i] is
// null where pass i runs the generic kernel.
constexpr uint32_t kJitInit = 1;     // the pass starts from |0...0> of each set instead of a load
constexpr uint32_t kJitAdjoint = 2;  // a reverse-sweep pass of the adjoint (records: the sweep's AGate)
constexpr uint32_t kJitWide = 4;     // 12-qubit tiles (records and TileSpec of the wide planner)
constexpr uint32_t kJitSweep = 8;    // an exp_value sweep of one string group (records: SweepKey)
constexpr uint32_t kJitPermStore = 16;  // Block swaps folded in: the tile is stored through a second TileSpec
// the compile-time part of one string of a sweep: x as tile positions, z on
// the register positions, ny mod 4
struct SweepKey {
    uint32_t xt, zr, q, pad;
};
struct JitPass {
    const void *gates;  // host tile::TGate records of the pass (one set)
    uint32_t count;     // 0: no specialised kernel
    uint32_t flags;
};
int32_t jit_prepare(int device, const std::vector<JitPass> &passes, std::vector<void *> &fns);
// the code object of one pass's kernel, compiled on the host (inspection, tests)
int32_t jit_code(const JitPass &p, std::vector<char> &code);
// set_shift < 32: per-set records (tile t takes gates + (t >> set_shift) * count);
// set_mask: the amplitude mask of one set (kJitInit passes)
// store_spec: the TileSpec the tile is stored through (kJitPermStore kernels)
hipError_t jit_launch(void *fn, double2 *amp, const void *spec, const void *gates, uint32_t ntiles,
                      hipStream_t stream, uint32_t set_shift = 32, uint64_t set_mask = ~0ull,
                      const void *store_spec = nullptr, uint32_t order = 3);
// an exp_value sweep (11-qubit tiles): strings = the group's device string
// records; one partial sum per workgroup into partial[0..grid)
hipError_t jit_launch_sweep(void *fn, const double2 *amp, const void *spec, const void *strings, uint32_t ntiles,
                            double *partial, uint32_t grid, hipStream_t stream);
// a wide pass (spec: the wide tile's tile::TileSpec of spec_bytes bytes)
// with the tile order `order` (tile_of_block in tile_dev.h), `threads` per
// workgroup (the wide tile's thread count)
hipError_t jit_launch_spec(void *fn, double2 *amp, const void *spec, size_t spec_bytes, const void *gates,
                           uint32_t ntiles, hipStream_t stream, uint32_t order, uint32_t threads);
// Tile-order tuner for compiled passes (tile_jit.hip).  Each pass layout
// (kernel, tile qubits, tile count) tries the candidate orders on its first
// runs, one per run, timed by events around its launch, and then keeps the
// fastest; the order only permutes which workgroup runs which tile.
constexpr uint32_t kDefaultTileOrder = 3;
struct OrderTrial {
    void *slot;         // the layout's tuning record (null: no trial this run)
    uint32_t order;     // the order to launch with
    hipEvent_t ev[2];   // around the launch when slot is set
};
// the order for this launch; a trial when the layout is still being tuned
// (the caller then records t.ev around the launch and reports the time)
OrderTrial order_pick(int device, void *fn, const void *spec, size_t spec_bytes, uint32_t ntiles);
// reports a trial after its launch completed (destroys its events)
void order_report(OrderTrial &t);
// Dependency-ordered runs in 12-qubit tiles with pass-specialised kernels
// (tiled_wide.hip, KBWH_RUN_COMPILED on flat states of 12 qubits and more).
// *done is false when the run was left to the 11-qubit passes (a pass kernel
// did not compile; nothing was launched).  A wide tile is kWideTileQubits
// qubits over 2^kWideTileTBits threads (6 lane bits, the rest wave bits), 16
// amplitudes per thread.
constexpr uint32_t kWideTileQubits = 12;
constexpr uint32_t kWideTileTBits = 8;
int32_t run_compiled_wide(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops, const double *params,
                          bool *done);
// the gate records (tile::TGate of the wide unit) of pass `pass` of the wide
// plan of ops on num_qubits qubits (host only; code inspection)
int32_t wide_pass_records(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits, uint32_t pass,
                          std::vector<unsigned char> &recs, uint32_t &count);
// a compiled reverse-sweep pass (the generic k_adj_tile's arguments; spec: the
// sweep's 12-qubit tile::TileSpec)
hipError_t jit_launch_adj(void *fn, double2 *amp, const void *spec, size_t spec_bytes, const void *gates, uint32_t tps,
                          uint32_t wps, double *partial, uint32_t grid, hipStream_t stream);
void free_adjoint(kbwh_state *s);
// Tiled reverse sweep of the adjoint gradient (n >= 11): slots[j] receives the
// term of the j-th parameterised op; partial holds kAdjTiledScratch doubles.
constexpr uint32_t kAdjTiledGrid = 1024;
constexpr uint32_t kAdjTiledSlots = 24;  // parameterised gates per pass
constexpr size_t kAdjTiledScratch = (size_t)kAdjTiledGrid * kAdjTiledSlots;
int32_t adjoint_reverse_tiled(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                              double *slots, double *partial, size_t partial_cap);
// The reverse sweep over nsets parameter sets (n >= 11): phi holds set k's
// 2^n amplitudes at k 2^n, lambda lies 2^lam_qubit amplitudes after phi;
// slots[k * (parameterised ops) + j] receives set k's term of the j-th
// parameterised op.  params_dev: nsets x num_params on the device.
int32_t adjoint_reverse_sets(kbwh_state *s, double2 *phi, uint32_t n, uint32_t nsets, uint32_t lam_qubit,
                             const kbwh_op *ops, uint64_t num_ops, const double *params_dev, uint32_t num_params,
                             double *slots);
int32_t queue_gate(kbwh_state *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask);
int32_t flush_queue(kbwh_state *s);
bool queue_pending(const kbwh_state *s);
int32_t flush_through_last_swap(kbwh_state *s);
// Physical events in program order (gates and Block swaps) in exact tiled
// passes, each swap folded into a pass when its qubits fit the tile (else a
// k_swap_bits pass of its own); one launch per op below 11 qubits.
int32_t run_events(kbwh_state *s, double2 *amp, const std::vector<kbwh_phys_event> &ev);
int32_t run_fused(kbwh_state *s, double2 *amp, const kbwh_op *ops, uint64_t num_ops, const double *params,
                  int k);

// Pauli strings flattened to masks: x = X|Y qubits, z = Z|Y qubits, ny = #Y.
struct PauliMasks {
    uint64_t x, z;
    uint32_t ny;
};
// Device memory an exp_value evaluation may use instead of the handle's own
// buffers (which grow with hipMalloc): a string table of strings_cap records
// of kExpStringBytes, block partials (kExpPartials doubles) and per-group or
// per-string results (kResultSlots doubles).
constexpr size_t kExpStringBytes = 24, kExpPartials = 4096 + 64;
struct ExpWork {
    void *strings;
    size_t strings_cap;
    double *partial;
    double *results;
};
int32_t exp_value_tiled(kbwh_state *s, const std::vector<PauliMasks> &ms, const double *coeffs, double *out,
                        const ExpWork *work = nullptr);
// Copies count doubles of device results (stream-ordered) to the host through
// the handle's pinned buffer.
int32_t reduce_from(kbwh_state *s, const double *src, double *out, uint32_t count);
// lam = sum_k coeffs[k] P_k phi in grouped tile sweeps.  n: the qubits the
// strings act on (0: the handle's); ntiles_total: tiles to sweep (0: all of
// an n-qubit state; parameter sets one after another have more).
int32_t pauli_apply_tiled(kbwh_state *s, const double2 *phi, double2 *lam, const std::vector<PauliMasks> &ms,
                          const double *coeffs, uint32_t n = 0, uint64_t ntiles_total = 0, uint32_t *sweeps = nullptr);
// Forward run of a circuit over nsets parameter sets laid out one after
// another in amp (set k's 2^n amplitudes at k 2^n): dependency-ordered tiled
// passes over the circuit's qubits (n >= 11), each set with its own angles
// params_dev[k * num_params + op.param] (a device array).
int32_t run_sets_forward(kbwh_state *s, double2 *amp, uint32_t n, uint32_t nsets, const kbwh_op *ops,
                         uint64_t num_ops, const double *params_dev, uint32_t num_params, kbwh_set_buffers &sb);
// Grows a device buffer to at least bytes (stream-ordered: waits for the
// handle's stream before freeing the old one).
int32_t grow_device(kbwh_state *s, void **buf, size_t *cap, size_t bytes);
int32_t pauli_masks(const kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                    uint32_t lo, uint32_t hi, bool physical, PauliMasks *out);

}  // namespace kbwh

// Apply queued live gates before the state is read or replaced.
#define FLUSH(s)                                        \
    do {                                                \
        const int32_t frc_ = kbwh::flush_queue(s);      \
        if (frc_) return frc_;                          \
    } while (0)
