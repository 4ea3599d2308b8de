/* Reductions and dump shared by the oracle backends.  TEST INFRASTRUCTURE. */
#include <stdlib.h>

#include "oracle_internal.h"

#define P1_CHUNK 4096u

/*
 * measure_p1 (dense_v2.rs:279-294, dense.rs:336-352): sum of
 * hypot(re, im)^2 over amplitudes whose bit `target` is 1.  Order: ascending
 * inside chunks of 4096 amplitudes, then chunk partials ascending.
 */
double orc_sum_p1(const orc_c64 *a, uint64_t size, uint32_t target) {
    const uint64_t mask = (uint64_t)1 << target;
    const uint64_t nchunk = (size + P1_CHUNK - 1) / P1_CHUNK;
    double *part = (double *)malloc(nchunk * sizeof(double));
#pragma omp parallel for schedule(static) if (nchunk > 16)
    for (uint64_t c = 0; c < nchunk; c++) {
        uint64_t lo = c * P1_CHUNK, hi = lo + P1_CHUNK;
        if (hi > size) hi = size;
        double acc = 0.0;
        for (uint64_t i = lo; i < hi; i++) {
            if ((i & mask) == mask) {
                double h = hypot(a[i].re, a[i].im);
                acc += h * h;
            }
        }
        part[c] = acc;
    }
    double total = 0.0;
    for (uint64_t c = 0; c < nchunk; c++) total += part[c];
    free(part);
    return total;
}

/* dump (dense_v2.rs:308-336): ascending index, keep |a| = hypot > 1e-15 */
uint64_t orc_dump_amps(const orc_c64 *a, uint64_t size, const uint32_t *qubits, uint32_t nq,
                       uint64_t *keys, double *re, double *im) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < size; i++) {
        if (hypot(a[i].re, a[i].im) > ORC_SMALL_EPSILON) {
            if (keys) {
                keys[m] = orc_dump_key(i, qubits, nq);
                re[m] = a[i].re;
                im[m] = a[i].im;
            }
            m++;
        }
    }
    return m;
}
