/*
 * Oracle restatement of Dense<f64> (src/dense.rs), the double-buffered twin of
 * DenseV2.  Non-diagonal gates gather each output amplitude from the current
 * buffer (dense.rs:91-309); diagonal gates act in place (dense.rs:153-169,
 * 211-229, 311-334).  Used only to cross-check the DenseV2 restatement.
 * TEST INFRASTRUCTURE.
 */
#include <math.h>
#include <stdlib.h>

#include "kbw_oracle.h"
#include "oracle_internal.h"

struct orc_v1 {
    uint32_t n;
    uint64_t size;
    orc_c64 *s0, *s1;
    int state; /* 1: current state lives in s0 (dense.rs:45-46) */
};

orc_v1 *orc_v1_new(uint32_t num_qubits) {
    if (num_qubits > 32) return NULL; /* dense.rs:71-73 */
    orc_v1 *s = (orc_v1 *)calloc(1, sizeof *s);
    s->n = num_qubits;
    s->size = (uint64_t)1 << num_qubits;
    s->s0 = (orc_c64 *)calloc(s->size, sizeof(orc_c64));
    s->s1 = (orc_c64 *)calloc(s->size, sizeof(orc_c64));
    s->s0[0].re = 1.0;
    s->state = 1;
    return s;
}

void orc_v1_free(orc_v1 *s) {
    if (!s) return;
    free(s->s0);
    free(s->s1);
    free(s);
}

/* dense.rs:60-66 */
orc_c64 *orc_v1_data(orc_v1 *s) { return s->state ? s->s0 : s->s1; }

/* dense.rs:51-58: flip the flag, return (current, next) */
static void get_states(orc_v1 *s, orc_c64 **cur, orc_c64 **next) {
    s->state = !s->state;
    if (s->state) {
        *cur = s->s1;
        *next = s->s0;
    } else {
        *cur = s->s0;
        *next = s->s1;
    }
}

/* num-complex `a * b` (Mul): re = re*br - im*bi; im = re*bi + im*br */
static inline orc_c64 cmul_v(orc_c64 a, orc_c64 b) {
    orc_c64 r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

void orc_v1_gate_mask(orc_v1 *s, int kind, double angle, uint32_t t, uint64_t m) {
    const uint64_t bit = (uint64_t)1 << t;
    const uint64_t n = s->size;
    if (kind == ORC_Z || kind == ORC_P || kind == ORC_RZ) {
        orc_c64 *a = orc_v1_data(s);
        orc_c64 ph0 = {0, 0}, ph1 = {0, 0};
        if (kind == ORC_P) ph1 = orc_cexp_i(angle); /* dense.rs:212 */
        if (kind == ORC_RZ) {                     /* dense.rs:312-313 */
            ph0 = orc_cexp_i(-angle / 2.0);
            ph1 = orc_cexp_i(angle / 2.0);
        }
#pragma omp parallel for schedule(static) if (n > 65536)
        for (uint64_t i = 0; i < n; i++) {
            if ((i & m) != m) continue;
            if (kind == ORC_Z) { /* dense.rs:165-167: amp = -amp */
                if (i & bit) {
                    a[i].re = -a[i].re;
                    a[i].im = -a[i].im;
                }
            } else if (kind == ORC_P) { /* dense.rs:225-227 */
                if (i & bit) a[i] = orc_cmul(a[i], ph1);
            } else { /* dense.rs:326-332 */
                a[i] = orc_cmul(a[i], (i & bit) ? ph1 : ph0);
            }
        }
        return;
    }
    orc_c64 *cur, *next;
    get_states(s, &cur, &next);
    const double c = cos(angle / 2.0), sn = sin(angle / 2.0);
    const double f = ORC_FRAC_1_SQRT_2;
#pragma omp parallel for schedule(static) if (n > 65536)
    for (uint64_t i = 0; i < n; i++) {
        if ((i & m) != m) {
            next[i] = cur[i];
            continue;
        }
        const orc_c64 cu = cur[i], fl = cur[i ^ bit];
        const int one = (i & bit) != 0;
        orc_c64 r;
        switch (kind) {
        case ORC_X: /* dense.rs:101,108-112 */
            r = fl;
            break;
        case ORC_Y: { /* dense.rs:127-132: flip * (one ? i : -i) */
            const orc_c64 pi = {0.0, 1.0}, mi = {-0.0, -1.0};
            r = cmul_v(fl, one ? pi : mi);
            break;
        }
        case ORC_H: { /* dense.rs:181-188: (cur * sign + flip) * FRAC_1_SQRT_2 */
            const double sign = one ? -1.0 : 1.0;
            r.re = (cu.re * sign + fl.re) * f;
            r.im = (cu.im * sign + fl.im) * f;
            break;
        }
        case ORC_RX: /* dense.rs:246-247 */
            r.re = c * cu.re + sn * fl.im;
            r.im = c * cu.im - sn * fl.re;
            break;
        case ORC_RY: { /* dense.rs:281-287 */
            const double sign = one ? 1.0 : -1.0;
            r.re = c * cu.re + sign * sn * fl.re;
            r.im = c * cu.im + sign * sn * fl.im;
            break;
        }
        default:
            r = cu;
        }
        next[i] = r;
    }
}

double orc_v1_measure_p1(orc_v1 *s, uint32_t target) {
    return orc_sum_p1(orc_v1_data(s), s->size, target);
}

/* dense.rs:354-367 */
void orc_v1_measure_collapse(orc_v1 *s, uint32_t target, int result, double p) {
    orc_c64 *cur, *next;
    get_states(s, &cur, &next);
    const uint64_t bit = (uint64_t)1 << target;
    for (uint64_t i = 0; i < s->size; i++) {
        if (((i & bit) != 0) == (result != 0)) {
            next[i].re = cur[i].re * p;
            next[i].im = cur[i].im * p;
        } else {
            next[i].re = 0.0;
            next[i].im = 0.0;
        }
    }
}
