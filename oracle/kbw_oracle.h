/*
 * CPU oracle for the dense state-vector path: a plain-C restatement of the
 * reference's DenseV2<f64> / Dense<f64> backends, the QuantumExecution trait
 * defaults, the QubitManager measure loop and the Block partitioner.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (kbw_amd/, libkbwh.so)
 * links, loads or calls this code.  It is used by tests/ (as the checker),
 * by __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline
 * leg (as the timed CPU baseline).
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * written operand-for-operand as the reference's num-complex / f64 code and
 * the library is compiled with -ffp-contract=off, so no multiply-add is
 * contracted (Rust never contracts a*b+c).  Trigonometry is glibc cos/sin,
 * the libm calls Rust's f64::cos / f64::sin make on Linux.
 *
 * Amplitudes are interleaved {re, im} doubles, index i = basis state i
 * (dense_v2.rs:45).
 */
#ifndef KBW_ORACLE_H
#define KBW_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } orc_c64;

/* Gate kinds; the same numbering as the product's C ABI. */
enum { ORC_X = 0, ORC_Y = 1, ORC_Z = 2, ORC_H = 3, ORC_P = 4, ORC_RX = 5, ORC_RY = 6, ORC_RZ = 7 };

/* Pauli letters for exp_value (quantum_execution.rs:114-121). */
enum { ORC_PAULI_X = 1, ORC_PAULI_Y = 2, ORC_PAULI_Z = 3 };

/* One circuit operation; identical layout to the product's kbwh_op. */
typedef struct {
    uint8_t kind;
    uint8_t target;
    uint16_t reserved;
    int32_t param;     /* parameter id, -1 when the angle is a constant */
    double angle;
    uint64_t ctrl_mask;
} orc_op;

/* ---- DenseV2<f64> (dense_v2.rs) ---------------------------------------- */
typedef struct orc_v2 orc_v2;
orc_v2 *orc_v2_new(uint32_t num_qubits);               /* NULL if > 32 (dense_v2.rs:155-157) */
void orc_v2_free(orc_v2 *s);
uint32_t orc_v2_num_qubits(const orc_v2 *s);
orc_c64 *orc_v2_data(orc_v2 *s);
void orc_v2_clear(orc_v2 *s);
void orc_v2_pauli_x(orc_v2 *s, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_pauli_y(orc_v2 *s, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_pauli_z(orc_v2 *s, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_hadamard(orc_v2 *s, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_phase(orc_v2 *s, double lambda, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_rx(orc_v2 *s, double theta, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_ry(orc_v2 *s, double theta, uint32_t target, const uint32_t *control, uint32_t nc);
void orc_v2_rz(orc_v2 *s, double theta, uint32_t target, const uint32_t *control, uint32_t nc);
/* kind-dispatched form with a control bit mask */
void orc_v2_gate_mask(orc_v2 *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask);
void orc_v2_run(orc_v2 *s, const orc_op *ops, uint64_t nops, const double *params);
double orc_v2_measure_p1(orc_v2 *s, uint32_t target);
void orc_v2_measure_collapse(orc_v2 *s, uint32_t target, int result, double p);
/* Two-call dump: keys==NULL returns the entry count only. */
uint64_t orc_v2_dump(orc_v2 *s, const uint32_t *qubits, uint32_t nq, uint64_t *keys, double *re,
                     double *im);

/* ---- Dense<f64> (dense.rs), the double-buffered twin ------------------- */
typedef struct orc_v1 orc_v1;
orc_v1 *orc_v1_new(uint32_t num_qubits);
void orc_v1_free(orc_v1 *s);
orc_c64 *orc_v1_data(orc_v1 *s);
void orc_v1_gate_mask(orc_v1 *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask);
double orc_v1_measure_p1(orc_v1 *s, uint32_t target);
void orc_v1_measure_collapse(orc_v1 *s, uint32_t target, int result, double p);

/* ---- QubitManager / trait defaults (quantum_execution.rs) -------------- */
/* rand's UniformFloat::new(0, total) scale and WeightedIndex pick. */
double orc_uniform_scale(double total);
uint64_t orc_measure(orc_v2 *s, const uint32_t *qubits, uint32_t nq, const double *uniforms);
double orc_exp_value(orc_v2 *s, const uint8_t *paulis, const uint32_t *qubits,
                     const uint32_t *offsets, const double *coeffs, uint32_t nstrings);
/* The same terms summed in 2^20-position chunks (parallel), then the chunk
 * partials in order: equal bits for n <= 20, rounding-close above. */
double orc_exp_value_chunked(orc_v2 *s, const uint8_t *paulis, const uint32_t *qubits,
                             const uint32_t *offsets, const double *coeffs, uint32_t nstrings);
/* Sampling with the faithful sequential cumulative sum over dump entries. */
int orc_sample_sequential(orc_v2 *s, const uint32_t *qubits, uint32_t nq, uint64_t shots,
                          const double *uniforms, uint64_t *outcomes);
/* Sampling with the hierarchical 64-way cumulative order shared with the GPU. */
int orc_sample_hier(const orc_c64 *amp, uint32_t num_qubits, const uint32_t *qubits, uint32_t nq,
                    uint64_t shots, const double *uniforms, uint64_t *outcomes);

/* ---- Block<(), DenseV2<f64>> (block.rs) -------------------------------- */
typedef struct orc_block orc_block;
orc_block *orc_block_new(uint32_t num_qubits, uint32_t block_size);  /* NULL if n <= 1 */
void orc_block_free(orc_block *b);
void orc_block_gate(orc_block *b, int kind, double angle, uint32_t target, const uint32_t *control,
                    uint32_t nc);
double orc_block_measure_p1(orc_block *b, uint32_t target);
void orc_block_measure_collapse(orc_block *b, uint32_t target, int result, double p);
uint64_t orc_block_dump(orc_block *b, const uint32_t *qubits, uint32_t nq, uint64_t *keys,
                        double *re, double *im);
/* logical amplitudes (flat order) after flushing the queue */
void orc_block_logical_state(orc_block *b, orc_c64 *out);
/* physical-order amplitudes (block-major) after flushing the queue */
void orc_block_physical_state(orc_block *b, orc_c64 *out);
void orc_block_map(const orc_block *b, uint32_t *logical_to_physical);
uint64_t orc_block_swap_count(const orc_block *b);
void orc_block_clear(orc_block *b);

#ifdef __cplusplus
}
#endif
#endif
