/*
 * Oracle restatement of DenseV2<f64> (src/dense_v2.rs).  TEST INFRASTRUCTURE:
 * used by tests/, smoke() and bench.py's cpu_baseline only.
 *
 * Parallel structure mirrors the reference's rayon use with OpenMP: the
 * outer loop runs over chunks of 2^(t+1) amplitudes (dense_v2.rs:93-100) and
 * the inner pair loop is split across threads only when the half-chunk has
 * at least 2048 pairs (dense_v2.rs:62,71).  Results do not depend on the
 * schedule: every amplitude pair is written by exactly one iteration.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kbw_oracle.h"
#include "oracle_internal.h"

struct orc_v2 {
    uint32_t n;
    uint64_t size;
    orc_c64 *a;
};

/* Gate coefficients the reference computes on the host before its loop. */
typedef struct {
    double c, s;        /* RX/RY: cos(theta/2), sin(theta/2) */
    orc_c64 ph0, ph1;   /* P: ph1 = exp(i*lambda); RZ: ph0 = exp(-i theta/2), ph1 = exp(i theta/2) */
} coef_t;

static coef_t make_coef(int kind, double angle) {
    coef_t k;
    memset(&k, 0, sizeof k);
    switch (kind) {
    case ORC_P:  /* dense_v2.rs:212  exp(i * lambda) */
        k.ph1 = orc_cexp_i(angle);
        break;
    case ORC_RX: /* dense_v2.rs:224-225 */
    case ORC_RY: /* dense_v2.rs:245-246 */
        k.c = cos(angle / 2.0);
        k.s = sin(angle / 2.0);
        break;
    case ORC_RZ: /* dense_v2.rs:266-267 */
        k.ph0 = orc_cexp_i(-angle / 2.0);
        k.ph1 = orc_cexp_i(angle / 2.0);
        break;
    default:
        break;
    }
    return k;
}

/* The per-pair closures of dense_v2.rs:166-277, operand for operand. */
static inline void pair_op(int kind, const coef_t *k, orc_c64 *ket0, orc_c64 *ket1) {
    orc_c64 t0 = *ket0, t1 = *ket1;
    switch (kind) {
    case ORC_X: /* dense_v2.rs:168-170  swap */
        *ket0 = t1;
        *ket1 = t0;
        break;
    case ORC_Y: { /* dense_v2.rs:178-182  swap; ket0 *= -i; ket1 *= i */
        const orc_c64 mi = {-0.0, -1.0}, pi = {0.0, 1.0};
        *ket0 = orc_cmul(t1, mi);
        *ket1 = orc_cmul(t0, pi);
        break;
    }
    case ORC_H: { /* dense_v2.rs:200-205 */
        const double f = ORC_FRAC_1_SQRT_2;
        ket0->re = (t0.re + t1.re) * f;
        ket0->im = (t0.im + t1.im) * f;
        ket1->re = (t0.re - t1.re) * f;
        ket1->im = (t0.im - t1.im) * f;
        break;
    }
    case ORC_RX: /* dense_v2.rs:229-237 */
        ket0->re = k->c * t0.re + k->s * t1.im;
        ket0->im = k->c * t0.im - k->s * t1.re;
        ket1->re = k->s * t0.im + k->c * t1.re;
        ket1->im = -k->s * t0.re + k->c * t1.im;
        break;
    case ORC_RY: /* dense_v2.rs:250-258 */
        ket0->re = k->c * t0.re - k->s * t1.re;
        ket0->im = k->c * t0.im - k->s * t1.im;
        ket1->re = k->s * t0.re + k->c * t1.re;
        ket1->im = k->s * t0.im + k->c * t1.im;
        break;
    case ORC_RZ: /* dense_v2.rs:271-272 */
        *ket0 = orc_cmul(t0, k->ph0);
        *ket1 = orc_cmul(t1, k->ph1);
        break;
    default:
        break;
    }
}

/* The per-amplitude closures of the diagonal gates (dense_v2.rs:190-191,215-216). */
static inline void diag_op(int kind, const coef_t *k, orc_c64 *ket1) {
    if (kind == ORC_Z) {
        const orc_c64 m1 = {-1.0, -0.0};
        *ket1 = orc_cmul(*ket1, m1);
    } else { /* ORC_P */
        *ket1 = orc_cmul(*ket1, k->ph1);
    }
}

/* dense_v2.rs:48-101 (pair engine) */
static void v2_gate(orc_v2 *s, int kind, const coef_t *k, uint32_t target, uint64_t ctrl_mask) {
    const uint64_t half = (uint64_t)1 << target;
    const uint64_t full = half << 1;
    const uint64_t low = ctrl_mask & (half - 1);
    const uint64_t high = ctrl_mask & ~(full - 1);
    const uint64_t nchunks = s->size / full;
    orc_c64 *a = s->a;
    if (half < 2048 || nchunks >= 64) {
#pragma omp parallel for schedule(static) if (nchunks > 1)
        for (uint64_t c = 0; c < nchunks; c++) {
            const uint64_t base = c * full;
            if ((base & high) != high) continue;
            for (uint64_t idx = 0; idx < half; idx++)
                if ((idx & low) == low) pair_op(kind, k, &a[base + idx], &a[base + half + idx]);
        }
    } else {
        for (uint64_t c = 0; c < nchunks; c++) {
            const uint64_t base = c * full;
            if ((base & high) != high) continue;
#pragma omp parallel for schedule(static)
            for (uint64_t idx = 0; idx < half; idx++)
                if ((idx & low) == low) pair_op(kind, k, &a[base + idx], &a[base + half + idx]);
        }
    }
}

/* dense_v2.rs:103-147 (diagonal engine: only the |1> half is touched) */
static void v2_diag(orc_v2 *s, int kind, const coef_t *k, uint32_t target, uint64_t ctrl_mask) {
    const uint64_t half = (uint64_t)1 << target;
    const uint64_t full = half << 1;
    const uint64_t low = ctrl_mask & (half - 1);
    const uint64_t high = ctrl_mask & ~(full - 1);
    const uint64_t nchunks = s->size / full;
    orc_c64 *a = s->a;
    if (half < 2048 || nchunks >= 64) {
#pragma omp parallel for schedule(static) if (nchunks > 1)
        for (uint64_t c = 0; c < nchunks; c++) {
            const uint64_t base = c * full;
            if ((base & high) != high) continue;
            for (uint64_t idx = 0; idx < half; idx++)
                if ((idx & low) == low) diag_op(kind, k, &a[base + half + idx]);
        }
    } else {
        for (uint64_t c = 0; c < nchunks; c++) {
            const uint64_t base = c * full;
            if ((base & high) != high) continue;
#pragma omp parallel for schedule(static)
            for (uint64_t idx = 0; idx < half; idx++)
                if ((idx & low) == low) diag_op(kind, k, &a[base + half + idx]);
        }
    }
}

/* dense_v2.rs:151-164 */
orc_v2 *orc_v2_new(uint32_t num_qubits) {
    if (num_qubits > 32) return NULL;
    orc_v2 *s = (orc_v2 *)calloc(1, sizeof *s);
    if (!s) return NULL;
    s->n = num_qubits;
    s->size = (uint64_t)1 << num_qubits;
    s->a = (orc_c64 *)malloc(s->size * sizeof(orc_c64));
    if (!s->a) {
        free(s);
        return NULL;
    }
    orc_v2_clear(s);
    return s;
}

void orc_v2_free(orc_v2 *s) {
    if (!s) return;
    free(s->a);
    free(s);
}

uint32_t orc_v2_num_qubits(const orc_v2 *s) { return s->n; }
orc_c64 *orc_v2_data(orc_v2 *s) { return s->a; }

/* dense_v2.rs:338-341 (first-touch is parallel so large states page in on all cores) */
void orc_v2_clear(orc_v2 *s) {
    orc_c64 *a = s->a;
#pragma omp parallel for schedule(static) if (s->size > 65536)
    for (uint64_t i = 0; i < s->size; i++) {
        a[i].re = 0.0;
        a[i].im = 0.0;
    }
    a[0].re = 1.0;
}

void orc_v2_gate_mask(orc_v2 *s, int kind, double angle, uint32_t target, uint64_t ctrl_mask) {
    coef_t k = make_coef(kind, angle);
    if (kind == ORC_Z || kind == ORC_P)
        v2_diag(s, kind, &k, target, ctrl_mask);
    else
        v2_gate(s, kind, &k, target, ctrl_mask);
}

static uint64_t mask_of(const uint32_t *control, uint32_t nc) {
    return orc_ctrl_mask(control, nc);
}

void orc_v2_pauli_x(orc_v2 *s, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_X, 0.0, t, mask_of(c, nc)); }
void orc_v2_pauli_y(orc_v2 *s, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_Y, 0.0, t, mask_of(c, nc)); }
void orc_v2_pauli_z(orc_v2 *s, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_Z, 0.0, t, mask_of(c, nc)); }
void orc_v2_hadamard(orc_v2 *s, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_H, 0.0, t, mask_of(c, nc)); }
void orc_v2_phase(orc_v2 *s, double l, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_P, l, t, mask_of(c, nc)); }
void orc_v2_rx(orc_v2 *s, double th, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_RX, th, t, mask_of(c, nc)); }
void orc_v2_ry(orc_v2 *s, double th, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_RY, th, t, mask_of(c, nc)); }
void orc_v2_rz(orc_v2 *s, double th, uint32_t t, const uint32_t *c, uint32_t nc) { orc_v2_gate_mask(s, ORC_RZ, th, t, mask_of(c, nc)); }

void orc_v2_run(orc_v2 *s, const orc_op *ops, uint64_t nops, const double *params) {
    for (uint64_t i = 0; i < nops; i++) {
        double angle = ops[i].angle;
        if (params && ops[i].param >= 0) angle = params[ops[i].param];
        orc_v2_gate_mask(s, ops[i].kind, angle, ops[i].target, ops[i].ctrl_mask);
    }
}

/*
 * dense_v2.rs:279-294: sum of norm()^2 = hypot(re, im)^2 over amplitudes with
 * bit `target` set.  The reference sums with rayon, whose tree depends on the
 * schedule; this restatement fixes the order: ascending within chunks of 4096
 * amplitudes, then the chunk partials in ascending order.
 */
double orc_v2_measure_p1(orc_v2 *s, uint32_t target) {
    return orc_sum_p1(s->a, s->size, target);
}

/* dense_v2.rs:296-306 */
void orc_v2_measure_collapse(orc_v2 *s, uint32_t target, int result, double p) {
    const uint64_t mask = (uint64_t)1 << target;
    orc_c64 *a = s->a;
#pragma omp parallel for schedule(static) if (s->size > 65536)
    for (uint64_t i = 0; i < s->size; i++) {
        if (((i & mask) == mask) == (result != 0)) {
            a[i].re = a[i].re * p;
            a[i].im = a[i].im * p;
        } else {
            a[i].re = 0.0;
            a[i].im = 0.0;
        }
    }
}

/* dense_v2.rs:308-336 */
uint64_t orc_v2_dump(orc_v2 *s, const uint32_t *qubits, uint32_t nq, uint64_t *keys, double *re,
                     double *im) {
    return orc_dump_amps(s->a, s->size, qubits, nq, keys, re, im);
}
