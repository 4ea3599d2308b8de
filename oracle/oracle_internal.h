/* Helpers shared by the oracle's translation units.  TEST INFRASTRUCTURE. */
#ifndef KBW_ORACLE_INTERNAL_H
#define KBW_ORACLE_INTERNAL_H

#include <math.h>
#include <stdint.h>

#include "kbw_oracle.h"

/* std::f64::consts::FRAC_1_SQRT_2 (used at dense_v2.rs:203-204) */
#define ORC_FRAC_1_SQRT_2 0.70710678118654752440
/* ket's small_epsilon for f64 (lib.rs:120-125), the dump threshold on |a| */
#define ORC_SMALL_EPSILON 1e-15

/* num-complex `a *= b` (MulAssign): re = re*br - im*bi; im = im*br + re*bi */
static inline orc_c64 orc_cmul(orc_c64 a, orc_c64 b) {
    orc_c64 r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.im * b.re + a.re * b.im;
    return r;
}

/* num-complex Complex::exp(i * x) = from_polar(exp(0*x), x) = (cos x, sin x) */
static inline orc_c64 orc_cexp_i(double x) {
    orc_c64 r;
    r.re = 1.0 * cos(x);
    r.im = 1.0 * sin(x);
    return r;
}

/* bitwise.rs:23-31 */
static inline uint64_t orc_ctrl_mask(const uint32_t *control, uint32_t nc) {
    uint64_t m = 0;
    for (uint32_t i = 0; i < nc; i++) m |= (uint64_t)1 << control[i];
    return m;
}

/* dump key packing (dense_v2.rs:315-321): the last listed qubit is bit 0 */
static inline uint64_t orc_dump_key(uint64_t state, const uint32_t *qubits, uint32_t nq) {
    uint64_t key = 0;
    for (uint32_t j = 0; j < nq; j++) {
        uint32_t q = qubits[nq - 1 - j];
        key |= ((state >> q) & 1u) << j;
    }
    return key;
}

double orc_sum_p1(const orc_c64 *a, uint64_t size, uint32_t target);
uint64_t orc_dump_amps(const orc_c64 *a, uint64_t size, const uint32_t *qubits, uint32_t nq,
                       uint64_t *keys, double *re, double *im);

#endif
