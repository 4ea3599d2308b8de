/*
 * kbwh.h -- C ABI of the MI355X-native dense state-vector engine.
 *
 * The engine is a drop-in for the dense path of KBW: the Rust trait
 * QuantumExecution (src/quantum_execution.rs:48-162) as implemented by
 * DenseV2<F> (src/dense_v2.rs:150-342) and Block<G, S>
 * (src/block.rs:211-381), which the C API selects by name in
 * kbw_make_configuration (src/c_api.rs:83-120, kbw.h:48-54).  Each entry
 * point below names the reference method it replaces.
 *
 * Conventions
 *  - Every function returns int32_t: KBWH_OK (0) on success, a KBWH_ERR_*
 *    code otherwise; kbwh_last_error() describes the last failure of the
 *    calling thread.  No C++ exception crosses this boundary.
 *  - Amplitudes are complex f64, interleaved {re, im}; basis index i holds
 *    |i>, qubit q is bit q of i (dense_v2.rs:45).
 *  - A handle owns its device memory and one HIP stream.  A handle is used by
 *    one thread at a time (the reference's &mut self); separate handles are
 *    independent.
 *  - Random draws are host-drawn uniforms u in [0, 1), one per draw, so the
 *    caller's RNG (the reference's StdRng seeded from KBW_SEED,
 *    quantum_execution.rs:182-193) stays on the caller's side.
 */
#ifndef KBWH_C_ABI_H_
#define KBWH_C_ABI_H_

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define KBWH_OK 0
#define KBWH_ERR_EXECUTION_FAILED 1   /* KetError::ExecutionFailed: n > 32 (dense_v2.rs:155-157), Block n <= 1 (block.rs:218-220) */
#define KBWH_ERR_INVALID_ARGUMENT 2   /* bad qubit index, null pointer, unknown gate kind */
#define KBWH_ERR_OUT_OF_MEMORY 3      /* device allocation failed */
#define KBWH_ERR_DEVICE 4             /* HIP runtime error (message in kbwh_last_error) */
#define KBWH_ERR_NO_DEVICE 5          /* no HIP device visible */
#define KBWH_ERR_ZERO_WEIGHTS 6       /* sample/measure over an all-zero distribution (the reference panics, convert.rs:60) */
#define KBWH_ERR_CAPACITY 7           /* caller buffer too small (kbwh_dump) */

/* ---- gate kinds (QuantumExecution methods, quantum_execution.rs:57-79) -- */
#define KBWH_X 0  /* pauli_x   dense_v2.rs:166-174 */
#define KBWH_Y 1  /* pauli_y   dense_v2.rs:176-186 */
#define KBWH_Z 2  /* pauli_z   dense_v2.rs:188-196 */
#define KBWH_H 3  /* hadamard  dense_v2.rs:198-209 */
#define KBWH_P 4  /* phase     dense_v2.rs:211-221 */
#define KBWH_RX 5 /* rx        dense_v2.rs:223-242 */
#define KBWH_RY 6 /* ry        dense_v2.rs:244-263 */
#define KBWH_RZ 7 /* rz        dense_v2.rs:265-277 */

/* Pauli letters of a Hamiltonian term (libket Pauli, quantum_execution.rs:114-121) */
#define KBWH_PAULI_X 1
#define KBWH_PAULI_Y 2
#define KBWH_PAULI_Z 3

/* kbwh_new flags */
#define KBWH_FLAG_BLOCK 1u  /* Block<_, DenseV2> semantics (block.rs): logical/physical map, LRU swaps */
#define KBWH_FLAG_QUEUE 2u  /* kbwh_gate queues gates; they are applied (same result bits) before any read of the
                               state: one launch for up to 12 qubits, tiled passes above.  With KBWH_FLAG_BLOCK the
                               queue is Block's gate queue (block.rs:179-208): the map, LRU stamps and swaps are
                               decided when a gate is given, gates and swaps run together in exact tiled passes
                               at the flush, and kbwh_measure_collapse does not flush (block.rs:311-316) unless
                               its target needs a swap.  kbwh_device_buffer users call kbwh_synchronize first. */

/* kbwh_run modes */
#define KBWH_RUN_PER_GATE 0u /* one HBM pass per gate: the plain path */
#define KBWH_RUN_TILED 1u    /* gates grouped into multi-gate passes over LDS tiles; same per-pair arithmetic, same result bits */
#define KBWH_RUN_FUSED 2u    /* gates fused into k-qubit blocks applied with FP64 MFMA; equal to rounding (~1e-16 per block), not bit for bit */
#define KBWH_RUN_AUTO 3u     /* the configuration chosen by kbwh_autotune for this handle (KBWH_RUN_TILED until tuned) */
#define KBWH_RUN_REORDER 4u  /* tiled passes over a dependency order of the gates (gates on disjoint qubits may trade places):
                                the same unitary, amplitudes equal to rounding; fewer passes.  On a flat state, uncontrolled
                                Z/P/RZ gates on one qubit are first fused across the gates diagonal on that qubit (phase
                                products).  A Block state keeps the reference's map and swap count (decided in program
                                order); its swaps are ordered with the gates */
#define KBWH_RUN_COMPILED 5u /* KBWH_RUN_REORDER's passes, each run by a kernel compiled for that pass (gate kinds,
                                tile positions and controls fixed in its code; angles read from the gate records):
                                the first run of a pass structure compiles its kernels at run time (hiprtc, in
                                parallel host threads) and later runs reuse them for the process; amplitudes equal to
                                KBWH_RUN_REORDER's to rounding.  From 12 qubits on the passes use 12-qubit tiles
                                (fewer passes) and the plan search (kbwh_set_plan_search); if one of those kernels
                                does not compile the run takes 11-qubit passes, where a pass that does not compile,
                                or of more than 256 gates, runs the generic kernel.  A Block state runs as
                                KBWH_RUN_REORDER */

/* One circuit operation (24 bytes).  `angle` is used unless `param` >= 0
 * and a parameter vector is supplied, in which case the angle is
 * params[param].  `ctrl_mask` has bit c set for each control qubit c. */
typedef struct kbwh_op {
    uint8_t kind;
    uint8_t target;
    uint16_t reserved;
    int32_t param;
    double angle;
    uint64_t ctrl_mask;
} kbwh_op;

typedef struct kbwh_state kbwh_state;

/* ---- lifetime ----------------------------------------------------------- */
/* Replaces DenseV2::new (dense_v2.rs:151-164) and Block::new (block.rs:212-248).
 * With KBWH_FLAG_BLOCK, block_size is the number of local qubits
 * (KBW_BLOCK_SIZE; 0 = read KBW_BLOCK_SIZE from the environment, default 20),
 * capped at n - 1 as in block.rs:222-228. */
int32_t kbwh_new(uint32_t num_qubits, uint32_t flags, uint32_t block_size, kbwh_state **out);
int32_t kbwh_free(kbwh_state *s);
/* Replaces clear (dense_v2.rs:338-341; block.rs:378-380 also resets the map). */
int32_t kbwh_clear(kbwh_state *s);
int32_t kbwh_num_qubits(const kbwh_state *s, uint32_t *n);

/* ---- gates ----------------------------------------------------------- */
/* Replaces pauli_x/pauli_y/pauli_z/hadamard/phase/rx/ry/rz
 * (quantum_execution.rs:57-79; dense_v2.rs:166-277).  A control equal to
 * the target is dropped, as the reference's split masks do
 * (dense_v2.rs:56-57).  `angle` is ignored by X/Y/Z/H. */
int32_t kbwh_gate(kbwh_state *s, int32_t kind, double angle, uint32_t target,
                  const uint32_t *control, uint32_t num_control);
int32_t kbwh_gate_mask(kbwh_state *s, int32_t kind, double angle, uint32_t target,
                       uint64_t ctrl_mask);
/* A gate sequence, as QubitManager::compute_gate replays a circuit
 * (quantum_execution.rs:310-314, 374-377).  `params` may be NULL. */
int32_t kbwh_run(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                 uint32_t mode);

/* Run configuration (no counterpart in the reference).  fusion_width: qubits
 * per fused block for KBWH_RUN_FUSED, 4 or 5 (default 5).  tile_targets:
 * most non-diagonal targets outside qubits 0..2 per tiled pass, 1..64
 * (default 64: as many as the tile's free positions, 8 for 11-qubit tiles). */
int32_t kbwh_set_run_config(kbwh_state *s, uint32_t fusion_width, uint32_t tile_targets);

/* enable != 0: the handle's adjoint gradients (kbwh_adjoint_gradient,
 * kbwh_adjoint_batch) run their forward and reverse passes on kernels compiled
 * for each pass, as KBWH_RUN_COMPILED, and kbwh_exp_value's grouped sweeps
 * (flat states of 11 qubits or more) run kernels compiled for each group of
 * strings (no counterpart in the reference; kbwh_exp_value_noalloc keeps the
 * generic sweep, since loading a kernel allocates).  The first call of a
 * circuit or string-group structure compiles them; values agree with the
 * generic kernels to rounding. */
int32_t kbwh_set_compiled(kbwh_state *s, int32_t enable);
/* Plan search of the dependency-ordered runs (KBWH_RUN_REORDER,
 * KBWH_RUN_COMPILED; no counterpart in the reference): besides the greedy
 * plan, `trials` randomised plans are tried and the one with the fewest passes
 * is kept, once per circuit structure (plans are cached for the process; the
 * angles do not change a plan).  trials < 0 restores the default: 0 for
 * KBWH_RUN_REORDER, 256 for KBWH_RUN_COMPILED (about 0.1 s of host time on the
 * first call of a 900-gate, 30-qubit circuit).  At most 65536. */
int32_t kbwh_set_plan_search(kbwh_state *s, int32_t trials);
/* Pass-specialised kernels of KBWH_RUN_COMPILED, for the process (no
 * counterpart in the reference): kernels compiled, passes that found theirs
 * already compiled, compile failures, passes run by the generic kernel and
 * the wall time spent compiling.  Any pointer may be NULL. */
int32_t kbwh_compiled_stats(uint64_t *kernels_compiled, uint64_t *cache_hits, uint64_t *failures,
                            uint64_t *generic_passes, double *compile_seconds);
/* Tile-order tuning of the compiled passes, for the process (no
 * counterpart in the reference).  The order in which workgroups take a
 * pass's tiles changes how its loads and stores fall on the HBM channels,
 * not its results (every order is a permutation of independent tiles).  Each
 * pass layout (kernel, tile qubits, tile count) of 4096 tiles or more (24
 * qubits in 12-qubit tiles, 23 in 11-qubit tiles) tries 4 candidate orders on
 * its first 4 runs, timed around its launch, and keeps the fastest.  Enabled by default; disabled, every pass
 * runs the default order. */
int32_t kbwh_set_order_tuning(int32_t enable);
/* Pass layouts whose order is chosen, layouts still trying candidates,
 * candidate launches timed so far, and chosen orders other than the default.
 * Any pointer may be NULL. */
int32_t kbwh_order_tuning_stats(uint64_t *layouts_tuned, uint64_t *layouts_pending, uint64_t *trials,
                                uint64_t *non_default);
/* The gfx950 code object of the specialised kernel of pass `pass` of the
 * KBWH_RUN_COMPILED plan of ops on a flat state of num_qubits (11..40)
 * qubits (12-qubit tiles from 12 qubits on, with the default plan search),
 * compiled on the host (no device needed; no counterpart in the reference).
 * With code NULL, *size receives the size in bytes. */
int32_t kbwh_compiled_pass_code(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits, uint32_t pass,
                                void *code, uint64_t *size);
/* The same for pass `pass` of the adjoint's reverse sweep of ops (the kernel
 * kbwh_set_compiled runs), on the host. */
int32_t kbwh_compiled_reverse_pass_code(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits, uint32_t pass,
                                        void *code, uint64_t *size);

/* The pass plan of KBWH_RUN_REORDER for a flat state of num_qubits qubits,
 * on the host only (no device needed; no counterpart in the reference):
 * order[0..num_ops) receives the op indices in execution order and
 * pass_end[0..*num_passes) the end of each pass in `order` (pass_end needs
 * room for num_ops entries).  tile_targets as in kbwh_set_run_config.
 * The plan covers the ops as given: a run fuses diagonal gates first. */
int32_t kbwh_reorder_plan(const kbwh_op *ops, uint64_t num_ops, uint32_t num_qubits, uint32_t tile_targets,
                          uint32_t *order, uint32_t *pass_end, uint64_t *num_passes);

/* Autotuner (no counterpart in the reference): times each candidate
 * configuration of `ops` (at most the first 256 gates) on a scratch buffer
 * the size of the state, leaving the state untouched, and keeps the fastest
 * for KBWH_RUN_AUTO.  Candidates: per-gate; tiled with all free positions,
 * 6 and 4 targets per pass; with allow_inexact, fused blocks of 4 and 5
 * qubits and dependency-ordered tiled passes.  ms[i] is the
 * time of candidate i in that order (0 when not run).  Needs device memory
 * for a second state (KBWH_ERR_OUT_OF_MEMORY otherwise). */
typedef struct kbwh_tune_result {
    uint32_t mode;          /* KBWH_RUN_PER_GATE, KBWH_RUN_TILED, KBWH_RUN_FUSED or KBWH_RUN_REORDER */
    uint32_t fusion_width;  /* for KBWH_RUN_FUSED */
    uint32_t tile_targets;  /* for KBWH_RUN_TILED and KBWH_RUN_REORDER (64: no cap) */
    uint32_t num_candidates;
    double ms[8];
} kbwh_tune_result;
int32_t kbwh_autotune(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                      int32_t allow_inexact, kbwh_tune_result *out);

/* ---- measurement ----------------------------------------------------- */
/* Replaces measure_p1 (dense_v2.rs:279-294). */
int32_t kbwh_measure_p1(kbwh_state *s, uint32_t target, double *p1);
/* Replaces measure_collapse (dense_v2.rs:296-306). */
int32_t kbwh_measure_collapse(kbwh_state *s, uint32_t target, int32_t result, double p);
/* Replaces QubitManager::do_measure (quantum_execution.rs:272-292): qubits
 * are measured in reverse list order, the k-th measured consumes
 * uniforms[k]; bit k of *result is its outcome. */
int32_t kbwh_measure(kbwh_state *s, const uint32_t *qubits, uint32_t num_qubits,
                     const double *uniforms, uint64_t *result);

/* ---- dump / sample / expectation -------------------------------------- */
/* Replaces dump (dense_v2.rs:308-336; block.rs:318-376).  Entries with
 * |a| > 1e-15 in physical index order; key bit j is qubit
 * qubits[num_qubits-1-j].  Two calls: keys == NULL stores the count only;
 * otherwise at most `capacity` entries are written (KBWH_ERR_CAPACITY if
 * more exist). */
int32_t kbwh_dump(kbwh_state *s, const uint32_t *qubits, uint32_t num_qubits, uint64_t capacity,
                  uint64_t *count, uint64_t *keys, double *re, double *im);
/* Replaces the default sample (quantum_execution.rs:99-102, convert.rs:54-73):
 * one dump key per shot, drawn with uniforms[shot] under the hierarchical
 * 64-way cumulative order (the WeightedIndex rule with a fixed association
 * of the additions). */
int32_t kbwh_sample(kbwh_state *s, const uint32_t *qubits, uint32_t num_qubits, uint64_t shots,
                    const double *uniforms, uint64_t *outcomes);
/* Replaces the default exp_value (quantum_execution.rs:109-158) with a direct
 * <psi|P|psi> reduction per Pauli string (the state is not modified).
 * String k is paulis/qubits[offsets[k] .. offsets[k+1]) with coefficient
 * coeffs[k]; *out = sum_k coeffs[k] * <P_k>. */
int32_t kbwh_exp_value(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                       const uint32_t *offsets, const double *coeffs, uint32_t num_strings,
                       double *out);

/* The same value as kbwh_exp_value, computed without allocating memory (no
 * counterpart in the reference).  For callers that cannot allocate: code
 * recorded into a HIP graph, tight serving loops, hosts that forbid device
 * allocation after start-up.  kbwh_exp_value may grow the handle's table
 * buffer with hipMalloc; this variant keeps its block partials, per-string
 * results and string table in one fixed device array per GPU (about 160 KiB,
 * held until the process ends).  Limits: at most 4096 strings
 * (KBWH_ERR_CAPACITY otherwise); no queued gates pending
 * (KBWH_ERR_INVALID_ARGUMENT: call kbwh_synchronize first, since a flush may
 * allocate).  All calls on one GPU share the array: calls that overlap in
 * time on one GPU (two threads, two handles) race and must be serialised by
 * the caller.  The result bits equal kbwh_exp_value's on a handle without
 * kbwh_set_compiled (with it, kbwh_exp_value's compiled sweeps agree to
 * rounding, this variant keeping the generic sweep). */
int32_t kbwh_exp_value_noalloc(kbwh_state *s, const uint8_t *paulis, const uint32_t *qubits,
                               const uint32_t *offsets, const double *coeffs, uint32_t num_strings,
                               double *out);

/* ---- adjoint differentiation (no counterpart in the reference: its
 * BatchExecution::gradient returns NativeGateUnsupported,
 * quantum_execution.rs:420-426, and libket falls back to parameter shift) --
 * Runs ops from |0...0> with angles params[op.param], then returns
 * E = <psi|H|psi> and dE/dparams[j] for every j < num_params, un-applying the
 * gates in reverse on two extra state vectors (no stored intermediates).
 * Parameterised ops must be RX, RY, RZ or P (any controls). */
int32_t kbwh_adjoint_gradient(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops,
                              const double *params, uint32_t num_params, const uint8_t *paulis,
                              const uint32_t *qubits, const uint32_t *offsets,
                              const double *coeffs, uint32_t num_strings, double *energy,
                              double *grad);

/* The same gradient for num_sets parameter sets of one circuit: set k's
 * angles are params[k * num_params + j]; energies[k] and grads[k * num_params
 * + j] receive its energy and gradient.  From 11 qubits the sets run in one
 * sweep: laid out one after another they form one state of
 * n + ceil(log2 num_sets) qubits whose top bits select the set, every tiled
 * pass covers all sets with each set's own gate records (coefficients
 * computed on the device, so values agree with kbwh_adjoint_gradient to
 * rounding), and energies and gradient terms are reduced per set.  Needs
 * 2^(n + ceil(log2 num_sets) + 1) amplitudes of device memory (phi and lambda
 * of every set).  Below 11 qubits the sets are run one by one.  At most 65535
 * sets. */
int32_t kbwh_adjoint_batch(kbwh_state *s, const kbwh_op *ops, uint64_t num_ops, const double *params,
                           uint32_t num_params, uint32_t num_sets, const uint8_t *paulis, const uint32_t *qubits,
                           const uint32_t *offsets, const double *coeffs, uint32_t num_strings, double *energies,
                           double *grads);

/* HBM passes of the last kbwh_adjoint_gradient / kbwh_adjoint_batch (no
 * counterpart in the reference): forward passes over phi, sweeps building
 * lambda = H phi (one read of phi and one read-modify-write of lambda each),
 * reverse passes (each reads and writes phi and lambda once). */
int32_t kbwh_adjoint_stats(const kbwh_state *s, uint64_t *forward_passes, uint64_t *lambda_sweeps,
                           uint64_t *reverse_passes);

/* ---- batches of circuits (no counterpart in the reference: its
 * BatchExecution replays one circuit per call on a fresh simulator,
 * quantum_execution.rs:367-427) --------------------------------------------
 * Many circuits of uneven qubit counts run in one persistent launch.  Circuit
 * c uses ops[first_op .. first_op+num_ops) (op.param indexes
 * params[first_param + op.param]) and the Pauli strings
 * [first_string .. first_string+num_strings) of the shared Hamiltonian
 * arrays (string k is paulis/qubits[offsets[k] .. offsets[k+1]) with weight
 * coeffs[k]).  kbwh_batch_run starts every circuit from |0...0> and writes
 * <H_c> for each circuit.  The gate structure is planned once at
 * kbwh_batch_new; only the parameters change between runs.  Angles of
 * parameterised gates are evaluated on the device, so results agree with the
 * per-gate path to ~1e-15, not bit for bit. */
typedef struct kbwh_circuit {
    uint32_t num_qubits;  /* 1..26 */
    uint32_t num_strings;
    uint64_t first_op;
    uint64_t num_ops;
    uint32_t first_string;
    uint32_t first_param;
} kbwh_circuit;

typedef struct kbwh_batch kbwh_batch;

int32_t kbwh_batch_new(const kbwh_circuit *circuits, uint32_t num_circuits, const kbwh_op *ops,
                       const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                       const double *coeffs, uint64_t num_params, kbwh_batch **out);
/* Plans the batch on the host only (no device needed) and checks the plan;
 * reports gate records, tiled passes and state amplitudes. */
int32_t kbwh_batch_plan(const kbwh_circuit *circuits, uint32_t num_circuits, const kbwh_op *ops,
                        const uint8_t *paulis, const uint32_t *qubits, const uint32_t *offsets,
                        const double *coeffs, uint64_t num_params, uint64_t *gates, uint64_t *passes,
                        uint64_t *amplitudes);
int32_t kbwh_batch_run(kbwh_batch *b, const double *params, double *exp_values);
/* enable != 0: the first run is captured in a hipGraph (parameter upload,
 * coefficient, reset and batch kernels, result download) and later runs
 * replay it. */
int32_t kbwh_batch_use_graph(kbwh_batch *b, int32_t enable);
/* The first `count` amplitudes of circuit `circuit` after the last run. */
int32_t kbwh_batch_state(kbwh_batch *b, uint32_t circuit, uint64_t count, double *out);
/* Gate records, tiled passes and amplitude-passes (sum over circuits of 2^n x passes; a group's
 * pass or read sweep counts once for all its sets, with 2^n x sets amplitudes). */
int32_t kbwh_batch_stats(const kbwh_batch *b, uint64_t *gates, uint64_t *passes, uint64_t *amplitudes);
/* Structure groups (no counterpart in the reference).  Circuits of a batch
 * whose ops (kinds, targets, controls, parameter slots, fixed angles) and
 * Pauli strings are the same, differing only in parameters and string
 * weights, run as one state of sets when the group spans at least 256 tiles
 * of 2^11 amplitudes: one launch per dependency-ordered pass over every set,
 * then read sweeps with per-set weights.  The other circuits run in the
 * persistent kernel.  A group is planned with at most 8 non-diagonal targets
 * above qubit 2 per pass and runs the generic pass kernel until configured. */
typedef struct kbwh_batch_group_info {
    uint32_t num_qubits;    /* state width of each set: max(n, 11) */
    uint32_t num_sets;      /* circuits in the group */
    uint32_t tile_targets;  /* target cap of its passes */
    uint32_t compiled;      /* passes run by pass-specialised kernels */
    uint32_t passes, sweeps;
    double tuned_ms;        /* the chosen configuration's time in the last kbwh_batch_autotune (0: not tuned) */
} kbwh_batch_group_info;
int32_t kbwh_batch_groups(const kbwh_batch *b, kbwh_batch_group_info *info, uint32_t capacity,
                          uint32_t *num_groups);
/* Every group: target cap 1..8 and, with compiled != 0, pass-specialised
 * kernels compiled at run time (as KBWH_RUN_COMPILED; a pass that does not
 * compile keeps the generic kernel).  Drops a captured graph. */
int32_t kbwh_batch_configure(kbwh_batch *b, uint32_t tile_targets, int32_t compiled);
/* Batch autotuner: per group, times its run with the generic kernel at
 * target caps 8, 6 and 4, then with compiled kernels at the best cap, and
 * keeps the fastest.  params: the parameter vector of the timing runs.
 * ms[4 g + i] receives candidate i's time for group g (capacity entries).
 * The batch's states are overwritten; drops a captured graph. */
int32_t kbwh_batch_autotune(kbwh_batch *b, const double *params, double *ms, uint32_t capacity);
int32_t kbwh_batch_free(kbwh_batch *b);

/* Fault injection (tests; no counterpart in the reference): after `after`
 * more device allocations by this library (any handle, any thread), the
 * next one fails as if the device were out of memory (the call that made it
 * returns KBWH_ERR_OUT_OF_MEMORY), and allocation works again afterwards.
 * A negative value cancels a pending failure. */
int32_t kbwh_inject_alloc_failure(int64_t after);

/* ---- raw access (tests, tooling) ------------------------------------- */
/* Copies amplitudes [first, first+count) of the physical buffer. */
int32_t kbwh_read_amplitudes(kbwh_state *s, uint64_t first, uint64_t count, double *out);
int32_t kbwh_write_amplitudes(kbwh_state *s, uint64_t first, uint64_t count, const double *in);
/* logical_to_physical[q] = physical bit of logical qubit q (identity unless Block). */
int32_t kbwh_qubit_map(const kbwh_state *s, uint32_t *logical_to_physical);
/* Block swaps made since the handle was created (kbwh_clear keeps the count). */
int32_t kbwh_swap_count(const kbwh_state *s, uint64_t *swaps);
/* Gates and HBM passes (kernel launches over the state) of the last kbwh_run. */
int32_t kbwh_run_stats(const kbwh_state *s, uint64_t *gates, uint64_t *passes);
int32_t kbwh_synchronize(kbwh_state *s);
/* The handle's HIP stream (hipStream_t) and device buffer, for event timing
 * and zero-copy views. */
int32_t kbwh_stream(const kbwh_state *s, void **stream);
int32_t kbwh_device_buffer(const kbwh_state *s, void **ptr, uint64_t *bytes);

/* ---- build info / errors --------------------------------------------- */
/* Mirrors kbw_build_info (c_api.rs:149-163, kbw.h:64). */
int32_t kbwh_build_info(const uint8_t **msg, uintptr_t *size);
const char *kbwh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
